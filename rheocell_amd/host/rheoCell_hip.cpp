// rheoCell_hip.cpp -- program entry with the HIP backend
// (mirrors src/rheoCell.cpp:13-49: input file plus key=value overrides,
// RCexception -> message and exit code 1).
#include <chrono>
#include <cstdio>

#include "rc_host.h"

int main(int argc, char **argv) {
	if(argc < 2) {
		fprintf(stderr, "Usage: %s input_file [key=value ...]\n", argv[0]);
		return 1;
	}
	try {
		input_file inp;
		inp.init_from_command_line_args(argc, argv);
		SimManager mysim(inp);
		mysim.load_options();
		mysim.init();
		const auto t0 = std::chrono::steady_clock::now();
		mysim.run();
		const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
		const rc_cg_stats_t cg = mysim.backend_ptr()->cg_stats();
		fprintf(stderr, "INFO: %lld steps of %d cells in %.3f s (%.4g cell-steps/s); CG solves %lld, iterations %lld\n",
				mysim.steps_run, mysim.backend_ptr()->N(), s, mysim.backend_ptr()->N() * (double) mysim.steps_run / s,
				cg.solves, cg.iters_total);
	}
	catch(RCexception &e) {
		fprintf(stderr, "ERROR: %s\n", e.what());
		return 1;
	}
	return 0;
}
