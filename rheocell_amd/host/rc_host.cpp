// rc_host.cpp -- host side of the HIP backend (see rc_host.h for the
// reference interfaces each class mirrors).
#include "rc_host.h"

#include <algorithm>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

// ---------------------------------------------------------------- errors
RCexception::RCexception(const char *fmt, ...) {
	char buf[2048];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	msg_ = buf;
}

// ---------------------------------------------------------------- input
static std::string trim(const std::string &s) {
	size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
	return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

void input_file::init_from_string(const std::string &text) {
	std::istringstream in(text);
	std::string line, pending_key, pending_val;
	int depth = 0;
	while(std::getline(in, line)) {
		if(depth > 0) {  // inside a { ... } block: keep the raw text
			for(char ch : line) {
				if(ch == '{') ++depth;
				if(ch == '}') --depth;
			}
			pending_val += "\n" + line;
			if(depth == 0) keys[pending_key] = trim(pending_val);
			continue;
		}
		const size_t hash = line.find('#');
		if(hash != std::string::npos) line = line.substr(0, hash);
		const size_t eq = line.find('=');
		if(eq == std::string::npos) continue;
		std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
		if(k.empty()) continue;
		for(char ch : v) {
			if(ch == '{') ++depth;
			if(ch == '}') --depth;
		}
		if(depth > 0) {
			pending_key = k;
			pending_val = v;
			continue;
		}
		keys[k] = v;
	}
}

void input_file::init_from_filename(const std::string &path) {
	std::ifstream f(path);
	if(!f.good()) throw RCexception("Can't read input file '%s'", path.c_str());
	std::stringstream ss;
	ss << f.rdbuf();
	init_from_string(ss.str());
}

void input_file::init_from_command_line_args(int argc, char **argv) {
	init_from_filename(argv[1]);
	for(int i = 2; i < argc; ++i) {
		std::string a = argv[i];
		const size_t eq = a.find('=');
		if(eq == std::string::npos) throw RCexception("Malformed command-line override '%s'", a.c_str());
		keys[trim(a.substr(0, eq))] = trim(a.substr(eq + 1));
	}
}

std::vector<std::string> input_file::unread_keys() const {
	std::vector<std::string> out;
	for(auto &kv : keys)
		if(!read.count(kv.first)) out.push_back(kv.first);
	return out;
}

static bool lookup(input_file *inp, const char *k, int mandatory, std::string &v) {
	auto it = inp->keys.find(k);
	if(it == inp->keys.end()) {
		if(mandatory) throw RCexception("Mandatory key `%s' not found", k);
		return false;
	}
	inp->read.insert(k);
	v = it->second;
	return true;
}

int getInputString(input_file *inp, const char *k, std::string &dest, int mandatory) {
	std::string v;
	if(!lookup(inp, k, mandatory, v)) return KEY_NOT_FOUND;
	dest = v;
	return KEY_FOUND;
}
int getInputNumber(input_file *inp, const char *k, number *dest, int mandatory) {
	std::string v;
	if(!lookup(inp, k, mandatory, v)) return KEY_NOT_FOUND;
	*dest = (number) std::atof(v.c_str());
	return KEY_FOUND;
}
int getInputInt(input_file *inp, const char *k, int *dest, int mandatory) {
	std::string v;
	if(!lookup(inp, k, mandatory, v)) return KEY_NOT_FOUND;
	*dest = (int) std::floor(std::atof(v.c_str()) + 0.1);
	return KEY_FOUND;
}
int getInputLLInt(input_file *inp, const char *k, llint *dest, int mandatory) {
	std::string v;
	if(!lookup(inp, k, mandatory, v)) return KEY_NOT_FOUND;
	*dest = (llint) std::floor(std::atof(v.c_str()) + 0.1);
	return KEY_FOUND;
}
int getInputBool(input_file *inp, const char *k, bool *dest, int mandatory) {
	std::string v;
	if(!lookup(inp, k, mandatory, v)) return KEY_NOT_FOUND;
	std::transform(v.begin(), v.end(), v.begin(), ::tolower);
	static const std::set<std::string> t = { "true", "1", "yes", "yup", "of course" };
	static const std::set<std::string> f = { "false", "0", "no", "nope", "are you crazy?" };
	if(t.count(v)) *dest = true;
	else if(f.count(v)) *dest = false;
	else throw RCexception("boolean key `%s' is invalid (`%s'), aborting.", k, v.c_str());
	return KEY_FOUND;
}
int getInputBoolAsInt(input_file *inp, const char *k, int *dest, int mandatory) {
	bool b;
	int r = getInputBool(inp, k, &b, mandatory);
	if(r == KEY_FOUND) *dest = (int) b;
	return r;
}

// ---------------------------------------------------------------- RNG
number Utils::gaussian() {
	static unsigned int isNextG = 0;
	static number nextG = -1e6;
	number toRet, u, v, w;
	if(isNextG) {
		isNextG = 0;
		return nextG;
	}
	w = 2.;
	while(w >= 1.0) {
		u = 2. * drand48() - 1.0;
		v = 2. * drand48() - 1.0;
		w = u * u + v * v;
	}
	w = std::sqrt((-2. * std::log(w)) / w);
	toRet = u * w;
	nextG = v * w;
	isNextG = 1;
	return toRet;
}

// ---------------------------------------------------------------- box / field
void SquareBox::init(int Lx, int Ly) {
	if(Lx != Ly) throw RCexception("The box in the configuration file is not a square");
	side = Lx;
}
int SquareBox::getElementX(int site, int distX) const {
	int x = (site - ((site / side) * side)) + distX;
	while(x < 0) x += side;
	while(x >= side) x -= side;
	return x;
}
int SquareBox::getElementY(int site, int distY) const {
	int y = (site / side) + distY;
	while(y < 0) y += side;
	while(y >= side) y -= side;
	return y;
}
int MultiPhaseField::GetSubIndex(int q, const SquareBox &box) const {
	return box.getElement(sub_corner_bottom_left, ((q - (q / LsubX) * LsubX) + offset[0]) % LsubX,
			((q / LsubX) + offset[1]) % LsubY);
}

// ---------------------------------------------------------------- interaction
void WetPolarModel::get_settings(input_file &inp) {
	std::string it = "wetpolarmodel";
	getInputString(&inp, "interaction_type", it, 0);
	std::transform(it.begin(), it.end(), it.begin(), ::tolower);
	if(it != "wetpolarmodel" && it != "wetmodel")
		throw RCexception("interaction_type `%s' is not implemented by the HIP backend", it.c_str());
	nematic = it == "wetmodel";
	if(nematic) {   // WetModel constructor values (WetModel.cpp:4-20)
		gamma = 0.01;
		mu = 3.;
	}
	getInputString(&inp, "topology", topology_filename, 1);
	getInputInt(&inp, "R", &R, 1);
	getInputNumber(&inp, "lambda", &lambda, 0);
	getInputNumber(&inp, "omega", &omega, 0);
	getInputNumber(&inp, "gamma", &gamma, 0);
	getInputNumber(&inp, "mu", &mu, 0);
	getInputNumber(&inp, "kappa", &kappa, 0);
	// the reference leaves these members uninitialised when absent
	// (WetPolarModel.cpp:4-21 vs :37-46): here they are mandatory
	getInputNumber(&inp, "zetaQ_self", &zetaQ_self_active, 1);
	getInputNumber(&inp, "zetaQ_inter", &zetaQ_inter_active, 1);
	getInputNumber(&inp, "J_Q", &J_Q_active, 1);
	getInputBool(&inp, "anchoring", &anchoring, 0);
	getInputNumber(&inp, "friction_cell", &friction_cell_active, 1);
	getInputNumber(&inp, "friction", &friction_active, 1);
	getInputNumber(&inp, "CGtolerance", &tolerance, 0);
	getInputNumber(&inp, "wall_slip", &wall_slip, 0);
	getInputNumber(&inp, "passive_alpha", &passive_alpha, 0);
	if(!nematic) getInputNumber(&inp, "beta", &beta_active, 1);   // WetModel has no beta
}

// getWalls of the reference's square wall boxes as their init builds it
// (SquareWalls.cpp:19-40, CircleWalls.cpp:19-37, TriangleWalls.cpp:19-50)
static std::vector<double> wall_field(const std::string &type, int L, number lambda_wall) {
	std::vector<double> w((size_t) L * L, 0.);
	if(type == "square_walls") {
		for(int y = 0; y < L; y++)
			for(int x = 0; x < L; x++)
				if(x < lambda_wall - 1 || x > L - lambda_wall || y < lambda_wall || y > L - lambda_wall) w[x + y * L] = 1.;
	}
	else if(type == "circle_walls") {
		for(int y = 0; y < L; y++)
			for(int x = 0; x < L; x++) {
				double dist = L / 2 - sqrt(((L / 2) - x) * ((L / 2) - x) + ((L / 2) - y) * ((L / 2) - y));
				w[x + y * L] = (dist <= 0 || dist < lambda_wall) ? 1. : 0.;
			}
	}
	else if(type == "triangle_walls") {
		std::fill(w.begin(), w.end(), 1.);
		int side = L - lambda_wall * 2;
		int height = int(round(side * sqrt(3.0) / 2.0));
		int center = L / 2, top = L / 2 - height / 2;
		for(int row = 0; row < height; ++row) {
			int half = int((side / 2.0) * (1.0 - (double) row / height));
			int y = top + row;
			if(y >= 0 && y < L)
				for(int x = center - half; x <= center + half && x < L; ++x)
					if(x >= 0) w[x + y * L] = 0.;
		}
	}
	return w;
}

// ForceFactory::make_forces (ForceFactory.cpp:61-150): the { key = value }
// blocks of external_forces_file, each a force on the fields it names (here
// every field, field = -1)
void FD_HIPBackend::read_external_forces(input_file &inp) {
	bool ext = false;
	getInputBool(&inp, "external_forces", &ext, 0);
	if(!ext) return;
	std::string file;
	getInputString(&inp, "external_forces_file", file, 1);
	bool json = false;
	getInputBool(&inp, "external_forces_as_JSON", &json, 0);
	if(json) throw RCexception("external_forces_as_JSON is not implemented by the HIP backend");
	std::ifstream in(file);
	if(!in.good()) throw RCexception("Can't read external_forces_file '%s'", file.c_str());
	std::string text, line;
	while(std::getline(in, line)) {
		const size_t c = line.find('#');
		text += (c == std::string::npos ? line : line.substr(0, c)) + "\n";
	}
	int depth = 0;
	size_t open = 0;
	for(size_t i = 0; i < text.size(); ++i) {
		if(text[i] == '{') {
			if(++depth > 1) throw RCexception("Syntax error in '%s': parentheses do not match", file.c_str());
			open = i + 1;
		}
		else if(text[i] == '}') {
			if(--depth < 0) throw RCexception("Syntax error in '%s': parentheses do not match", file.c_str());
			input_file f;
			f.init_from_string(text.substr(open, i - open));
			std::string type, field;
			getInputString(&f, "type", type, 1);
			getInputString(&f, "field", field, 1);
			if(field != "-1") throw RCexception("external force `%s': only field = -1 is implemented by the HIP backend", type.c_str());
			rc_ext_force F{};
			F.lambda = 2.;
			if(type == "channel_walls" || type == "moving_walls" || type == "circle_walls" || type == "triangle_walls" ||
					type == "square_walls") {
				F.kind = RC_FORCE_CHANNEL_WALLS;
				getInputNumber(&f, "kappa_wall", &F.kappa_wall, 0);
				getInputNumber(&f, "lambda", &F.lambda, 0);
			}
			else if(type == "shear_flow_channel" || type == "pois_flow_channel") {
				F.kind = type == "shear_flow_channel" ? RC_FORCE_SHEAR_FLOW : RC_FORCE_POIS_FLOW;
				getInputNumber(&f, "kappa_wall", &F.kappa_wall, 1);
				getInputNumber(&f, "lambda", &F.lambda, 1);
				getInputNumber(&f, "shear_rate", &F.shear_rate, 1);
			}
			else if(type == "stretching") {
				F.kind = RC_FORCE_STRETCHING;
				getInputNumber(&f, "shear_rate", &F.shear_rate, 1);
			}
			else throw RCexception("Invalid force type `%s'", type.c_str());
			if(forces.size() == RC_MAX_FORCES) throw RCexception("at most %d external forces", RC_MAX_FORCES);
			forces.push_back(F);
		}
	}
}

int WetPolarModel::get_N_from_topology() const {
	std::ifstream t(topology_filename);
	if(!t.good()) throw RCexception("Can't read topology file '%s'. Aborting", topology_filename.c_str());
	int n;
	t >> n;
	return n;
}

rc_model_params WetPolarModel::to_params(number J0, number dt) const {
	rc_model_params p{};
	p.gamma = gamma;
	p.lambda = lambda;
	p.omega = omega;
	p.mu = mu;
	p.kappa = kappa;
	p.R = R;
	p.anchoring = anchoring ? 1 : 0;
	p.zetaQ_self = zetaQ_self_active;
	p.zetaQ_inter = zetaQ_inter_active;
	p.J_Q = J_Q_active;
	p.friction = friction_active;
	p.friction_cell = friction_cell_active;
	p.CGtolerance = tolerance;
	p.wall_slip = wall_slip;
	p.passive_alpha = passive_alpha;
	p.beta = beta_active;
	p.J0 = J0;
	p.dt = dt;
	p.cg_max_iter = 0;
	p.interaction = nematic ? RC_INTERACTION_WETNEMATIC : RC_INTERACTION_WETPOLAR;
	return p;
}

// ---------------------------------------------------------------- backend
FD_HIPBackend::FD_HIPBackend() {}
FD_HIPBackend::~FD_HIPBackend() {
	writer.reset();   // writes what is still queued
	if(h) rc_destroy(h);
}

void FD_HIPBackend::check(int rc) {
	if(rc != RC_OK) throw RCexception("HIP backend: %s", h ? rc_last_error(h) : "(no handle)");
}

void FD_HIPBackend::get_settings(input_file &inp) {
	std::string b;
	getInputString(&inp, "backend", b, 1);
	if(b != "HIP" && b != "CUDA")
		throw RCexception("backend `%s' is not provided by this binary (it implements backend = HIP)", b.c_str());
	std::string st = "FD_CPU";
	getInputString(&inp, "sim_type", st, 0);
	if(st != "FD_CPU") throw RCexception("sim_type `%s' is not implemented by the HIP backend", st.c_str());
	std::string prec = "double";
	getInputString(&inp, "backend_precision", prec, 0);
	if(prec != "double") throw RCexception("the HIP backend computes in double precision only");
	std::string box = "square";
	getInputString(&inp, "box", box, 0);
	if(box != "square") throw RCexception("box `%s' is not implemented by the HIP backend", box.c_str());
	// BoxFactory::make_box (BoxFactory.cpp:16-26): square, or one of its wall boxes
	getInputString(&inp, "type", box_type, 0);
	if(box_type == "square_walls" || box_type == "circle_walls" || box_type == "triangle_walls")
		getInputNumber(&inp, "lambda_wall", &lambda_wall, 1);
	else box_type = "none";
	read_external_forces(inp);
	interaction.get_settings(inp);
	getInputBool(&inp, "restart_step_counter", &restart_step_counter, 0);
	std::string reload;
	if(getInputString(&inp, "reload_from", reload, 0) == KEY_FOUND) {
		std::string tmp;
		if(getInputString(&inp, "conf_file", tmp, 0) == KEY_FOUND)
			throw RCexception("Input file error: \"conf_file\" cannot be specified if \"reload_from\" is specified");
		if(restart_step_counter)
			throw RCexception("Input file error: \"restart_step_counter\" must be set to false if \"reload_from\" is specified");
		conf_filename = reload;
	}
	else
		getInputString(&inp, "conf_file", conf_filename, 1);
	std::string T;
	getInputString(&inp, "T", T, 1);
	getInputString(&inp, "trajectory_file", trajectory_file, 1);
	// configuration lines formatted and written by a writer thread (default), or inline
	getInputBool(&inp, "async_conf_output", &async_conf_output, 0);
	getInputString(&inp, "lastconf_file", lastconf_file, 0);
	getInputString(&inp, "equilibratedconf_file", equilibratedconf_file, 0);
	// ObservableOutput prefixes every output file name with output_prefix
	// (ObservableOutput.cpp:20-27)
	std::string prefix;
	getInputString(&inp, "output_prefix", prefix, 0);
	trajectory_file = prefix + trajectory_file;
	lastconf_file = prefix + lastconf_file;
	equilibratedconf_file = prefix + equilibratedconf_file;
	getInputNumber(&inp, "dt", &dt, 1);
	getInputNumber(&inp, "J0", &J0, 0);
	std::string wm;
	if(getInputString(&inp, "cg_warm_start", wm, 0) == KEY_FOUND) warm_mode = wm == "previous" ? RC_WARM_PREVIOUS : RC_WARM_REFERENCE;
}

// SimBackend::read_next_configuration (SimBackend.cpp:220-314) for one frame
void FD_HIPBackend::read_next_configuration() {
	std::ifstream in(conf_filename);
	if(!in.good()) throw RCexception("Can't read configuration file '%s'", conf_filename.c_str());
	std::string line;
	std::getline(in, line);
	llint step;
	if(sscanf(line.c_str(), "t = %lld", &step) != 1)
		throw RCexception("Malformed headers found in an input configuration.\"t = <int>\" was expected, but \"%s\" was found instead.", line.c_str());
	std::getline(in, line);
	double Lx, Ly;
	if(sscanf(line.c_str(), "b = %lf %lf", &Lx, &Ly) != 2)
		throw RCexception("Malformed headers found in an input configuration.\"b = <float> <float> <float>\" was expected, but \"%s\" was found instead.", line.c_str());
	box.init((int) Lx, (int) Ly);
	int i = 0;
	while(i < N() && std::getline(in, line)) {
		std::istringstream ss(line);
		std::vector<double> v;
		double d;
		while(ss >> d) v.push_back(d);
		if(v.size() < 11) break;
		MultiPhaseField &p = *fields[i];
		p.LsubX = (int) v[0];
		p.LsubY = (int) v[1];
		p.subSize = p.LsubX * p.LsubY;
		if(v.size() < 11 + 2 * (size_t) p.subSize) throw RCexception("configuration line of cell %d is truncated", i);
		p.thetaQ = p.thetaQ_old = PI * (1 - 2 * drand48());   // MultiPhaseField::init draw
		{   // the director of that draw (MultiPhaseField.cpp:80-95), kept by the reader
			const number q00 = cos(2 * p.thetaQ), q01 = sin(2 * p.thetaQ);
			const number nemQ_mod = sqrt(q00 * q00 + q01 * q01);
			const number nx = sqrt((1 + q00 / nemQ_mod) / 2);
			const number sgn = q01 > 0 ? 1 : q01 < 0 ? -1 : 0;
			const number ny = sgn * sqrt((1 - q00 / nemQ_mod) / 2);
			p.nemQ = { nx, ny };
			p.nemQ_old = { nx, ny };
		}
		p.x_sub_left = p.LsubX;
		p.y_sub_bottom = p.LsubY;
		p.area = p.perimeter = p.sumF = 0;
		p.S00 = p.S01 = 0;
		p.CoM = { v[2], v[3] };
		p.CoM_old = { v[2], v[3] };
		p.offset = { (int) v[4], (int) v[5] };
		p.sub_corner_bottom_left = p.sub_corner_bottom_left_old = (int) v[6];
		p.unrap_sub_corner_bottom_left_x = (int) v[7];
		p.unrap_sub_corner_bottom_left_y = (int) v[8];
		p.Q00 = v[9];
		p.Q01 = v[10];
		p.thetaQ = 0.5 * atan2(2 * p.Q01, 2 * p.Q00);
		p.thetaQ_old = p.thetaQ;
		p.fieldScalar.resize(p.subSize);
		p.map_sub_to_box.resize(p.subSize);
		for(int k = 0; k < p.subSize; ++k) {
			p.map_sub_to_box[k] = (int) v[11 + 2 * k];
			p.fieldScalar[k] = v[11 + 2 * k + 1];
			p.area += v[11 + 2 * k + 1] * v[11 + 2 * k + 1];
			p.sumF += v[11 + 2 * k + 1];
		}
		++i;
	}
	if(i != N())
		throw RCexception("The number of lines found in configuration file (%d) doesn't match the parsed number of particles (%d)", i, N());
	start_step_from_file = restart_step_counter ? 0 : step;
	curr_step = start_step_from_file;
}

void FD_HIPBackend::init() {
	// the trajectory is opened at start-up, appended to unless
	// restart_step_counter is set, which starts it afresh (ObservableOutput.cpp:31-33, 94-99)
	if(restart_step_counter) {
		std::ofstream t(trajectory_file, std::ios::trunc);
		if(!t) throw RCexception("Stream %s not writable", trajectory_file.c_str());
	}
	const int n = interaction.get_N_from_topology();
	fields.clear();
	for(int i = 0; i < n; ++i) {
		fields.emplace_back(new MultiPhaseField());
		fields.back()->index = i;
		// run_clock draw (WetPolarModel.cpp:72); WetModel::read_topology draws nothing
		if(!interaction.nematic) (void) (int) (2 * drand48() * interaction.J_Q_active);
	}
	read_next_configuration();
	rc_box_desc d{};
	d.L = box.side;
	d.ncells = n;
	d.params = interaction.to_params(J0, dt);
	if(box_type != "none") {
		walls = wall_field(box_type, box.side, lambda_wall);
		d.walls = walls.data();
	}
	d.nforces = (int) forces.size();
	for(size_t f = 0; f < forces.size(); ++f) d.forces[f] = forces[f];
	int rc = rc_create(&d, 1, &h);
	if(rc != RC_OK) {
		std::string msg = h ? rc_last_error(h) : "rc_create failed";
		if(h) rc_destroy(h);
		h = nullptr;
		throw RCexception("HIP backend: %s", msg.c_str());
	}
	check(rc_set_warm_start(h, warm_mode));
	apply_changes_to_simulation_data();
	check(rc_init_forces(h));   // FD_CPUBackend::init force evaluation
	host_stale = true;
}

void FD_HIPBackend::apply_changes_to_simulation_data() {
	std::vector<rc_cell> cells(N());
	std::vector<double> phi;
	interaction.field_start_index.assign(N(), 0);
	int start = 0;
	for(int i = 0; i < N(); ++i) {
		const MultiPhaseField &p = *fields[i];
		rc_cell &c = cells[i];
		c.LsubX = p.LsubX;
		c.LsubY = p.LsubY;
		c.offset[0] = p.offset[0];
		c.offset[1] = p.offset[1];
		c.corner = p.sub_corner_bottom_left;
		c.corner_old = p.sub_corner_bottom_left_old;
		c.x_sub_left = p.x_sub_left;
		c.y_sub_bottom = p.y_sub_bottom;
		c.CoM[0] = p.CoM[0];
		c.CoM[1] = p.CoM[1];
		c.CoM_old[0] = p.CoM_old[0];
		c.CoM_old[1] = p.CoM_old[1];
		c.unrap[0] = p.unrap_sub_corner_bottom_left_x;
		c.unrap[1] = p.unrap_sub_corner_bottom_left_y;
		c.Q00 = p.Q00;
		c.Q01 = p.Q01;
		c.thetaQ = p.thetaQ;
		c.thetaQ_old = p.thetaQ_old;
		c.area = p.area;
		c.area_old = p.area_old;
		c.sumF = p.sumF;
		c.sumF_old = p.sumF_old;
		c.perimeter = p.perimeter;
		c.S00 = p.S00;
		c.S01 = p.S01;
		c.total_force_repulsion[0] = p.total_force_repulsion_x;
		c.total_force_repulsion[1] = p.total_force_repulsion_y;
		c.nemQ[0] = p.nemQ[0];
		c.nemQ[1] = p.nemQ[1];
		c.nemQ_old[0] = p.nemQ_old[0];
		c.nemQ_old[1] = p.nemQ_old[1];
		phi.insert(phi.end(), p.fieldScalar.begin(), p.fieldScalar.end());
		interaction.field_start_index[i] = start;
		start += p.subSize;
	}
	check(rc_upload_cells(h, cells.data(), phi.data()));
}

void FD_HIPBackend::apply_simulation_changes_after_equilibration() { check(rc_apply_changes_after_equilibration(h)); }

void FD_HIPBackend::sim_step() { sim_steps(1); }

void FD_HIPBackend::sim_steps(int n) {
	if(n <= 0) return;
	if(interaction.nematic) {   // WetModel draws no noise
		check(rc_step(h, n, nullptr));
		host_stale = true;
		return;
	}
	std::vector<double> g((size_t) n * N());
	for(auto &x : g) x = Utils::gaussian();   // predictor draws, cell order (WetPolarModel.cpp:365)
	check(rc_step(h, n, g.data()));
	host_stale = true;
}

void FD_HIPBackend::apply_simulation_data_changes() {
	if(!host_stale) return;
	std::vector<rc_cell> cells(N());
	long long rows = 0;
	check(rc_total_rows(h, &rows));
	std::vector<double> phi(rows);
	check(rc_download_cells(h, cells.data(), phi.data()));
	interaction.vec_v_x.resize(rows);
	interaction.vec_v_y.resize(rows);
	check(rc_get_rows(h, 0, interaction.vec_v_x.data()));
	check(rc_get_rows(h, 1, interaction.vec_v_y.data()));
	size_t start = 0;
	for(int i = 0; i < N(); ++i) {
		MultiPhaseField &p = *fields[i];
		const rc_cell &c = cells[i];
		p.LsubX = c.LsubX;
		p.LsubY = c.LsubY;
		p.subSize = c.LsubX * c.LsubY;
		p.offset = { c.offset[0], c.offset[1] };
		p.sub_corner_bottom_left = c.corner;
		p.sub_corner_bottom_left_old = c.corner_old;
		p.x_sub_left = c.x_sub_left;
		p.y_sub_bottom = c.y_sub_bottom;
		p.CoM = { c.CoM[0], c.CoM[1] };
		p.CoM_old = { c.CoM_old[0], c.CoM_old[1] };
		p.unrap_sub_corner_bottom_left_x = c.unrap[0];
		p.unrap_sub_corner_bottom_left_y = c.unrap[1];
		p.Q00 = c.Q00;
		p.Q01 = c.Q01;
		p.thetaQ = c.thetaQ;
		p.thetaQ_old = c.thetaQ_old;
		p.area = c.area;
		p.area_old = c.area_old;
		p.sumF = c.sumF;
		p.sumF_old = c.sumF_old;
		p.perimeter = c.perimeter;
		p.S00 = c.S00;
		p.S01 = c.S01;
		p.total_force_repulsion_x = c.total_force_repulsion[0];
		p.total_force_repulsion_y = c.total_force_repulsion[1];
		p.nemQ = { c.nemQ[0], c.nemQ[1] };
		p.nemQ_old = { c.nemQ_old[0], c.nemQ_old[1] };
		p.fieldScalar.assign(phi.begin() + start, phi.begin() + start + p.subSize);
		p.map_sub_to_box.resize(p.subSize);
		for(int q = 0; q < p.subSize; ++q) p.map_sub_to_box[q] = p.GetSubIndex(q, box);
		interaction.field_start_index[i] = (int) start;
		start += p.subSize;
	}
	host_stale = false;
}

// Configuration::headers / field (src/Observables/Configuration.cpp:36-80)
std::string format_conf(const ConfSnapshot &s) {
	std::stringstream conf;
	conf.precision(15);
	conf << "t = " << s.step << std::endl;
	conf << "b = " << s.Lx << " " << s.Ly << std::endl;
	for(size_t i = 0; i < s.cells.size(); ++i) {
		const ConfSnapshot::Cell &p = s.cells[i];
		if(i) conf << std::endl;
		conf << p.LsubX << " " << p.LsubY << " " << p.CoM[0] << " " << p.CoM[1] << " " << p.offset[0] << " "
			 << p.offset[1] << " " << p.corner << " " << p.unrap[0] << " " << p.unrap[1] << " " << p.Q00 << " "
			 << p.Q01 << " ";
		for(size_t q = 0; q < p.fieldScalar.size(); q++) conf << p.map_sub_to_box[q] << " " << p.fieldScalar[q] << " ";
	}
	return conf.str();
}

ConfSnapshot FD_HIPBackend::conf_snapshot(llint step) {
	ConfSnapshot s;
	s.step = step;
	s.Lx = box.getXsize();
	s.Ly = box.getYsize();
	s.cells.resize(N());
	for(int i = 0; i < N(); ++i) {
		const MultiPhaseField *p = fields[i].get();
		ConfSnapshot::Cell &c = s.cells[i];
		c.LsubX = p->LsubX;
		c.LsubY = p->LsubY;
		c.offset[0] = p->offset[0];
		c.offset[1] = p->offset[1];
		c.corner = p->sub_corner_bottom_left;
		c.CoM[0] = p->CoM[0];
		c.CoM[1] = p->CoM[1];
		c.unrap[0] = p->unrap_sub_corner_bottom_left_x;
		c.unrap[1] = p->unrap_sub_corner_bottom_left_y;
		c.Q00 = p->Q00;
		c.Q01 = p->Q01;
		c.map_sub_to_box.assign(p->map_sub_to_box.begin(), p->map_sub_to_box.begin() + p->subSize);
		c.fieldScalar.assign(p->fieldScalar.begin(), p->fieldScalar.begin() + p->subSize);
	}
	return s;
}

std::string FD_HIPBackend::conf_string(llint step) { return format_conf(conf_snapshot(step)); }

static void write_conf_files(const ConfSnapshot &s) {
	const std::string text = format_conf(s) + "\n";
	if(s.to_trajectory) {
		std::ofstream t(s.trajectory_file, std::ios::app);
		t << text;
		if(!t) throw RCexception("Stream %s not writable", s.trajectory_file.c_str());
	}
	std::ofstream l(s.lastconf_file, std::ios::trunc);
	l << text;
	if(!l) throw RCexception("Stream %s not writable", s.lastconf_file.c_str());
}

ConfWriter::ConfWriter() : th([this] { run(); }) {}

ConfWriter::~ConfWriter() {
	{
		std::lock_guard<std::mutex> g(m);
		quit = true;
	}
	cv.notify_all();
	th.join();
	// a write error nobody collected (the run left through another exception)
	if(!error.empty()) fprintf(stderr, "ERROR: configuration output: %s\n", error.c_str());
}

void ConfWriter::push(ConfSnapshot &&s) {
	std::unique_lock<std::mutex> g(m);
	cv.wait(g, [this] { return q.size() < 2 || !error.empty(); });
	if(!error.empty()) throw RCexception("%s", error.c_str());
	q.push_back(std::move(s));
	cv.notify_all();
}

void ConfWriter::drain() {
	std::unique_lock<std::mutex> g(m);
	cv.wait(g, [this] { return (q.empty() && !busy) || !error.empty(); });
	if(!error.empty()) throw RCexception("%s", error.c_str());
}

void ConfWriter::run() {
	std::unique_lock<std::mutex> g(m);
	for(;;) {
		cv.wait(g, [this] { return quit || !q.empty(); });
		if(!error.empty()) q.clear();   // after a write error nothing more is written (no gaps)
		if(q.empty()) {
			if(quit) return;   // quit with nothing left
			cv.notify_all();
			continue;
		}
		ConfSnapshot s = std::move(q.front());
		q.pop_front();
		busy = true;
		g.unlock();
		std::string err;
		try {
			write_conf_files(s);
		} catch(const std::exception &e) {
			err = e.what();
		}
		g.lock();
		busy = false;
		if(!err.empty() && error.empty()) error = err;
		cv.notify_all();
	}
}

void FD_HIPBackend::print_conf(bool only_last) {
	apply_simulation_data_changes();
	ConfSnapshot s = conf_snapshot(curr_step);
	s.to_trajectory = !only_last;
	s.trajectory_file = trajectory_file;
	s.lastconf_file = lastconf_file;
	if(!async_conf_output) {
		write_conf_files(s);
		return;
	}
	if(!writer) writer.reset(new ConfWriter());
	writer->push(std::move(s));
}

void FD_HIPBackend::drain_output() {
	if(writer) writer->drain();
}

void FD_HIPBackend::print_equilibration_info() {
	apply_simulation_data_changes();
	std::ofstream e(equilibratedconf_file, std::ios::trunc);
	e << conf_string(curr_step) << "\n";
}

rc_cg_stats_t FD_HIPBackend::cg_stats() {
	rc_cg_stats_t s{};
	check(rc_cg_stats(h, &s));
	return s;
}

// ---------------------------------------------------------------- manager
bool SimManager::stop = false;

static void gbl_terminate(int) { SimManager::stop = true; }

void SimManager::load_options() {
	getInputLLInt(&input, "steps", &steps, 1);
	getInputLLInt(&input, "equilibration_steps", &equilibration_steps, 0);
	bool rsc = false;
	getInputBool(&input, "restart_step_counter", &rsc, 0);
	if(!rsc && equilibration_steps > 0)
		throw RCexception("Incompatible key values found:\n\tif equilibration_steps > 0, restart_step_counter must be set to true.\n\tAborting");
	if(equilibration_steps < 0) throw RCexception("Equilibration steps can not be < 0. Aborting");
	if(getInputInt(&input, "seed", &seed, 0) == KEY_NOT_FOUND) seed = (int) time(nullptr);
}

void SimManager::init() {
	srand48(seed);
	backend.reset(new FD_HIPBackend());
	backend->get_settings(input);
	signal(SIGTERM, gbl_terminate);
	signal(SIGINT, gbl_terminate);
	signal(SIGUSR2, gbl_terminate);
	std::string ts = "linear";
	getInputString(&input, "time_scale", ts, 1);
	if(ts != "linear") throw RCexception("Time scale '%s' not supported by this build", ts.c_str());
	getInputInt(&input, "print_conf_interval", &print_conf_interval, 1);
	if(print_conf_interval <= 0) throw RCexception("print_conf_interval must be > 0");
	backend->init();
}

// SimManager::run (SimManager.cpp:131-178) with the linear time scale; the
// steps between two prints are handed to the device in one call.
void SimManager::run() {
	FD_HIPBackend &b = *backend;
	for(llint s = 0; s < equilibration_steps && !stop;) {
		const int n = (int) std::min<llint>(equilibration_steps - s, 256);
		b.sim_steps(n);
		s += n;
	}
	b.print_equilibration_info();
	b.apply_simulation_changes_after_equilibration();
	const llint start = b.start_step_from_file;
	llint next = start + print_conf_interval - (start % print_conf_interval);
	for(steps_run = 0; steps_run < steps && !stop;) {
		if(b.current_step() == next) {
			if(steps_run > 0) b.print_conf();
			next += print_conf_interval;
		}
		const llint until = std::min<llint>(steps - steps_run, next - b.current_step());
		const int n = (int) std::max<llint>(1, std::min<llint>(until, 256));
		b.sim_steps(n);
		for(int k = 0; k < n; ++k) b.increment_current_step();
		steps_run += n;
	}
	if(b.current_step() == next && steps_run > 0) b.print_conf();
	b.print_conf(true);
	b.drain_output();
}
