// rc_host.h -- C++ host side of the HIP backend, mirroring the reference's
// Backend / Field / Interaction / Box interfaces for the WetPolarModel path.
//
//   input_file + getInput*      src/Utilities/parse_input/parse_input.h:133-154
//   RCexception                 src/Utilities/RCexception.h
//   Utils::gaussian             src/Utilities/Utils.h:210-234 (glibc drand48)
//   SquareBox                   src/Boxes/SquareBox.cpp:21-82
//   wall boxes / external forces src/Boxes/SquareWalls.cpp, CircleWalls.cpp, TriangleWalls.cpp,
//                               src/Forces/ForceFactory.cpp:61-150
//   MultiPhaseField (host copy) src/Fields/BaseField.h:122-225, MultiPhaseField.cpp:348-359
//   WetPolarModel (parameters)  src/Interactions/WetPolarModel.cpp:4-93
//   FD_HIPBackend               src/Backends/SimBackend.h:86-158, FD_CPUBackend.cpp:13-160
//   SimManager                  src/Managers/SimManager.cpp:53-178
//
// All simulation arithmetic runs on the GPU behind include/rheocell.h; this
// layer reads inputs and confs, owns the random stream, and writes confs.
#pragma once

#include <condition_variable>
#include <cstdarg>
#include <deque>
#include <map>
#include <mutex>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rheocell.h"

typedef double number;
typedef long long int llint;

#define KEY_FOUND 0
#define KEY_NOT_FOUND -1
#define PI 3.141592653589793238462643f

class RCexception : public std::exception {
	std::string msg_;

public:
	RCexception(const char *fmt, ...);
	const char *what() const noexcept override { return msg_.c_str(); }
};

// key = value input file, '#' comments, '{ ... }' blocks kept as raw text,
// command-line key=value overrides (parse_input.cpp:20-156)
struct input_file {
	std::map<std::string, std::string> keys;
	std::set<std::string> read;
	void init_from_filename(const std::string &path);
	void init_from_string(const std::string &text);
	void init_from_command_line_args(int argc, char **argv);  // argv[1] = file, argv[2..] = overrides
	std::vector<std::string> unread_keys() const;
};

int getInputString(input_file *inp, const char *skey, std::string &dest, int mandatory);
int getInputNumber(input_file *inp, const char *skey, number *dest, int mandatory);
int getInputInt(input_file *inp, const char *skey, int *dest, int mandatory);
int getInputLLInt(input_file *inp, const char *skey, llint *dest, int mandatory);
int getInputBool(input_file *inp, const char *skey, bool *dest, int mandatory);
int getInputBoolAsInt(input_file *inp, const char *skey, int *dest, int mandatory);

namespace Utils {
number gaussian();  // Marsaglia polar method on drand48, cached second variate
}

class SquareBox {
public:
	int side = -1;
	void init(int Lx, int Ly);
	int getElementX(int site, int distX) const;
	int getElementY(int site, int distY) const;
	int getElement(int site, int distX, int distY) const { return getElementX(site, distX) + getElementY(site, distY) * side; }
	int getXsize() const { return side; }
	int getYsize() const { return side; }
};

// Host copy of a cell (the members the observables read)
struct MultiPhaseField {
	int index = -1;
	int LsubX = 30, LsubY = 30, subSize = 900;
	std::vector<number> CoM{ 0., 0. }, CoM_old{ 0., 0. };
	std::vector<int> offset{ 0, 0 };
	int sub_corner_bottom_left = 0, sub_corner_bottom_left_old = 0;
	number unrap_sub_corner_bottom_left_x = 0, unrap_sub_corner_bottom_left_y = 0;
	int x_sub_left = 0, y_sub_bottom = 0;
	number Q00 = 0, Q01 = 0, thetaQ = 0, thetaQ_old = 0;
	number area = 0, area_old = 0, sumF = 0, sumF_old = 0, perimeter = 0, S00 = 0, S01 = 0;
	number total_force_repulsion_x = 0, total_force_repulsion_y = 0;
	std::vector<number> nemQ{ 0., 0. }, nemQ_old{ 0., 0. };   // nematic director (WetModel)
	std::vector<number> fieldScalar;
	std::vector<int> map_sub_to_box;
	int GetSubIndex(int q, const SquareBox &box) const;
};

// The interaction's parameters and the velocity rows (WetPolarModel.h:101-102,
// WetModel.h:99-100): interaction_type = wetpolarmodel, or the nematic
// wetmodel (nematic set), whose keys are the same less beta
class WetPolarModel {
public:
	bool nematic = false;
	number gamma = 0.06, lambda = 2.5, omega = 0.004, mu = 20., kappa = 0.1;
	int R = 8;
	bool anchoring = false;
	number zetaQ_self_active = 0, zetaQ_inter_active = 0, J_Q_active = 0;
	number friction_active = 0, friction_cell_active = 0, beta_active = 0;
	number tolerance = 0.0001, wall_slip = 0.5, passive_alpha = 1.;
	bool have_active[6] = { false, false, false, false, false, false };
	std::string topology_filename;
	std::vector<number> vec_v_x, vec_v_y;
	std::vector<int> field_start_index;
	void get_settings(input_file &inp);
	int get_N_from_topology() const;
	number get_velocity_x(const MultiPhaseField *p, int q) const { return vec_v_x[q + field_start_index[p->index]]; }
	number get_velocity_y(const MultiPhaseField *p, int q) const { return vec_v_y[q + field_start_index[p->index]]; }
	rc_model_params to_params(number J0, number dt) const;
};

// What Configuration::headers / field print of one step (Configuration.cpp:36-80),
// copied off the live fields so that it can be formatted while steps go on.
struct ConfSnapshot {
	llint step = 0;
	int Lx = 0, Ly = 0;
	struct Cell {
		int LsubX, LsubY, offset[2], corner;
		number CoM[2], unrap[2], Q00, Q01;
		std::vector<int> map_sub_to_box;
		std::vector<number> fieldScalar;
	};
	std::vector<Cell> cells;
	bool to_trajectory = true;       // append to the trajectory (not only the last conf)
	std::string trajectory_file, lastconf_file;
};
std::string format_conf(const ConfSnapshot &s);   // the text Configuration writes, precision 15

// Configuration output off the step loop (no counterpart in the reference,
// which formats on its one thread): print_conf hands a snapshot to one writer
// thread, which formats it and appends it in print order.  At most two
// snapshots wait; a further print waits for the writer.
class ConfWriter {
public:
	ConfWriter();
	~ConfWriter();
	void push(ConfSnapshot &&s);
	void drain();   // every pushed snapshot written; rethrows a write error
private:
	void run();
	std::mutex m;
	std::condition_variable cv;
	std::deque<ConfSnapshot> q;
	bool quit = false, busy = false;
	std::string error;
	std::thread th;
};

class FD_HIPBackend {
public:
	FD_HIPBackend();
	~FD_HIPBackend();
	void get_settings(input_file &inp);
	void init();
	void sim_step();                      // one Heun step (SimBackend::sim_step)
	void sim_steps(int n);                // n steps in one call, same random stream
	void apply_simulation_data_changes();  // device -> host copies
	void apply_changes_to_simulation_data();  // host -> device
	void apply_simulation_changes_after_equilibration();
	void print_conf(bool only_last = false);
	void print_equilibration_info();
	std::string conf_string(llint step);
	ConfSnapshot conf_snapshot(llint step);
	void drain_output();   // waits for the configuration writer
	llint current_step() const { return curr_step; }
	void increment_current_step() { ++curr_step; }
	int N() const { return (int) fields.size(); }
	rc_cg_stats_t cg_stats();

	std::vector<std::unique_ptr<MultiPhaseField>> fields;
	WetPolarModel interaction;
	SquareBox box;
	llint start_step_from_file = 0;
	llint curr_step = 0;
	number dt = 0, J0 = 1.0;

private:
	void read_next_configuration();
	void read_external_forces(input_file &inp);
	void check(int rc);
	rc_handle *h = nullptr;
	std::string conf_filename, trajectory_file, lastconf_file = "last_conf.dat",
			equilibratedconf_file = "equilibrated_conf.dat";
	bool restart_step_counter = false;
	bool host_stale = true;
	int warm_mode = RC_WARM_REFERENCE;
	bool async_conf_output = true;   // input key async_conf_output (no counterpart in the reference)
	std::string box_type = "none";   // type: none or a square wall box
	number lambda_wall = 0;
	std::vector<double> walls;        // getWalls of the box (empty: none)
	std::vector<rc_ext_force> forces; // external forces (external_forces_file)
	std::unique_ptr<ConfWriter> writer;
};

class SimManager {
public:
	explicit SimManager(input_file &inp) : input(inp) {}
	void load_options();
	void init();
	void run();
	static bool stop;
	FD_HIPBackend *backend_ptr() { return backend.get(); }
	llint steps = 0, equilibration_steps = 0, steps_run = 0;
	int seed = -1;
	int print_conf_interval = 0;

private:
	input_file &input;
	std::unique_ptr<FD_HIPBackend> backend;
};
