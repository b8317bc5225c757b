// ref_driver.cpp -- golden-vector dumper linked against the reference's own
// MultiPhaseField / SquareBox / Utils sources (compiled from /root/reference
// by oracle/build_ref.sh; the output binary lives in oracle/_ref/ only).
//
// Reads commands from stdin, one per line, writes one JSON object per line:
//   trig L                          -> MultiPhaseField cos/sin tables (resizing)
//   rng seed n                      -> n drand48 values after srand48(seed)
//   gauss seed n                    -> n Utils::gaussian() values after srand48(seed)
//   subindex Lbox LX LY offx offy corner
//                                   -> GetSubIndex for every storage site
//   nbrsub LX LY                    -> neighbors_sub table (periodic in the patch)
//   setpos Lbox LX LY offx offy corner ux uy comx comy xsl ysb phi_0 ... phi_{S-1}
//                                   -> MultiPhaseField::set_positions(box) result
//   stencil Lbox LX LY offx offy corner phi_0 ... phi_{S-1}
//                                   -> BaseInteraction::derivX / derivY / Laplacian
//                                      (BaseInteraction.h:196-263) at every site
//   walls type Lbox lambda_wall     -> getWalls of every box site of the wall box
//                                      type (square_walls, circle_walls, triangle_walls)
//   stencilw type Lbox lambda_wall wall_slip LX LY offx offy corner phi_0 ... phi_{S-1}
//                                   -> the stencils in that wall box
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "Boxes/CircleWalls.h"
#include "Boxes/SquareBox.h"
#include "Boxes/SquareWalls.h"
#include "Boxes/TriangleWalls.h"
#include <memory>
#include "Fields/MultiPhaseField.h"
#include "Interactions/BaseInteraction.h"
#include "Utilities/Utils.h"

// The reference's stencils are members of the abstract BaseInteraction; this
// driver-side subclass only makes it instantiable and exposes them.
class StencilProbe : public BaseInteraction {
public:
	void init() override {}
	void allocate_fields(std::vector<BaseField *> &) override {}
	void check_input_sanity(std::vector<BaseField *> &) override {}
	void set_wall_slip(double w) { wall_slip = w; }
	double dx(BaseField *p, int q, int k) { return derivX(p, q, k); }
	double dy(BaseField *p, int q, int k) { return derivY(p, q, k); }
	double lap(BaseField *p, int q, int k) { return Laplacian(p, q, k); }
};

// a wall box of the reference, set up as BoxFactory + get_settings + init do
static std::unique_ptr<BaseBox> wall_box(const std::string &type, int L, double lambda_wall) {
	std::unique_ptr<BaseBox> b;
	if(type == "square_walls") b.reset(new SquareWalls());
	else if(type == "circle_walls") b.reset(new CircleWalls());
	else b.reset(new TriangleWalls());
	input_file inp;
	inp.init_from_string("lambda_wall = " + std::to_string(lambda_wall));
	b->get_settings(inp);
	b->init(L, L);
	return b;
}

static void dump_vec(const char *name, const std::vector<double> &v) {
	printf("\"%s\": [", name);
	for(size_t i = 0; i < v.size(); ++i) printf("%s%.17g", i ? ", " : "", v[i]);
	printf("]");
}
static void dump_ivec(const char *name, const std::vector<int> &v) {
	printf("\"%s\": [", name);
	for(size_t i = 0; i < v.size(); ++i) printf("%s%d", i ? ", " : "", v[i]);
	printf("]");
}

int main() {
	std::string line;
	while(std::getline(std::cin, line)) {
		std::istringstream in(line);
		std::string cmd;
		if(!(in >> cmd)) continue;
		if(cmd == "trig") {
			int L;
			in >> L;
			MultiPhaseField f;
			f.init(L, L);
			printf("{\"cmd\": \"trig\", \"L\": %d, ", L);
			dump_vec("cos", f.cos_x_table);
			printf(", ");
			dump_vec("sin", f.sin_x_table);
			printf("}\n");
		}
		else if(cmd == "rng" || cmd == "gauss") {
			long seed;
			int n;
			in >> seed >> n;
			srand48(seed);
			std::vector<double> v;
			for(int i = 0; i < n; ++i) v.push_back(cmd == "rng" ? drand48() : Utils::gaussian());
			printf("{\"cmd\": \"%s\", \"seed\": %ld, ", cmd.c_str(), seed);
			dump_vec("values", v);
			printf("}\n");
		}
		else if(cmd == "subindex") {
			int Lb, LX, LY, ox, oy, corner;
			in >> Lb >> LX >> LY >> ox >> oy >> corner;
			SquareBox box;
			box.init(Lb, Lb);
			MultiPhaseField f;
			f.init(LX, LY);
			f.set_positions(ox, oy, corner, 0, 0, Lb);
			std::vector<int> k;
			for(int q = 0; q < f.subSize; ++q) k.push_back(f.GetSubIndex(q, &box));
			printf("{\"cmd\": \"subindex\", \"args\": [%d, %d, %d, %d, %d, %d], ", Lb, LX, LY, ox, oy, corner);
			dump_ivec("k", k);
			printf("}\n");
		}
		else if(cmd == "nbrsub") {
			int LX, LY;
			in >> LX >> LY;
			MultiPhaseField f;
			f.init(LX, LY);
			printf("{\"cmd\": \"nbrsub\", \"LX\": %d, \"LY\": %d, ", LX, LY);
			dump_ivec("nbr", f.neighbors_sub);
			printf("}\n");
		}
		else if(cmd == "setpos") {
			int Lb, LX, LY, ox, oy, corner, xsl, ysb;
			double ux, uy, cx, cy;
			in >> Lb >> LX >> LY >> ox >> oy >> corner >> ux >> uy >> cx >> cy >> xsl >> ysb;
			SquareBox box;
			box.init(Lb, Lb);
			MultiPhaseField f;
			f.init(LX, LY);
			f.set_positions(ox, oy, corner, 0, 0, Lb);
			f.unrap_sub_corner_bottom_left_x = ux;
			f.unrap_sub_corner_bottom_left_y = uy;
			f.CoM[0] = cx;
			f.CoM[1] = cy;
			f.x_sub_left = xsl;
			f.y_sub_bottom = ysb;
			for(int q = 0; q < f.subSize; ++q) in >> f.fieldScalar[q];
			f.set_positions(&box);
			printf("{\"cmd\": \"setpos\", \"LX\": %d, \"LY\": %d, \"offset\": [%d, %d], \"corner\": %d, "
				   "\"corner_old\": %d, \"unrap\": [%.17g, %.17g], \"CoM\": [%.17g, %.17g], \"area\": %.17g, ",
				   f.LsubX, f.LsubY, f.offset[0], f.offset[1], f.sub_corner_bottom_left, f.sub_corner_bottom_left_old,
				   f.unrap_sub_corner_bottom_left_x, f.unrap_sub_corner_bottom_left_y, f.CoM[0], f.CoM[1], f.area);
			dump_vec("phi", f.fieldScalar);
			printf("}\n");
		}
		else if(cmd == "walls") {
			std::string type;
			int Lb;
			double lw;
			in >> type >> Lb >> lw;
			auto box = wall_box(type, Lb, lw);
			std::vector<double> w;
			for(int k = 0; k < Lb * Lb; ++k) w.push_back(box->getWalls(k));
			printf("{\"cmd\": \"walls\", \"type\": \"%s\", \"L\": %d, \"lambda_wall\": %.17g, ", type.c_str(), Lb, lw);
			dump_vec("walls", w);
			printf("}\n");
		}
		else if(cmd == "stencil" || cmd == "stencilw") {
			std::string type = "none";
			double lw = 0, ws = 0.5;
			if(cmd == "stencilw") in >> type;
			int Lb, LX, LY, ox, oy, corner;
			in >> Lb;
			if(cmd == "stencilw") in >> lw >> ws;
			in >> LX >> LY >> ox >> oy >> corner;
			std::unique_ptr<BaseBox> boxp;
			if(cmd == "stencilw") boxp = wall_box(type, Lb, lw);
			else {
				boxp.reset(new SquareBox());
				boxp->init(Lb, Lb);
			}
			BaseBox &box = *boxp;
			MultiPhaseField f;
			f.init(LX, LY);
			f.set_positions(ox, oy, corner, 0, 0, Lb);
			for(int q = 0; q < f.subSize; ++q) in >> f.fieldScalar[q];
			StencilProbe probe;
			probe.set_box(&box);
			probe.set_wall_slip(ws);
			std::vector<double> dx, dy, lap;
			for(int q = 0; q < f.subSize; ++q) {
				int k = f.GetSubIndex(q, &box);
				dx.push_back(probe.dx(&f, q, k));
				dy.push_back(probe.dy(&f, q, k));
				lap.push_back(probe.lap(&f, q, k));
			}
			printf("{\"cmd\": \"%s\", \"type\": \"%s\", \"lambda_wall\": %.17g, \"wall_slip\": %.17g, "
				   "\"args\": [%d, %d, %d, %d, %d, %d], ", cmd.c_str(), type.c_str(), lw, ws, Lb, LX, LY, ox, oy, corner);
			dump_vec("dx", dx);
			printf(", ");
			dump_vec("dy", dy);
			printf(", ");
			dump_vec("lap", lap);
			printf("}\n");
		}
		fflush(stdout);
	}
	return 0;
}
