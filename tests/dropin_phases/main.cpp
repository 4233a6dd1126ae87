// A user program written against fvhyper's API that calls the phases of one
// RK stage itself (explicit.h:88-149) instead of run(): the drop-in's phase
// functions must give the oracle's phases bit for bit.  The plugin is the
// forward-step physics (Roe flux in the joined spelling, Venkatakrishnan
// limiter, mirror walls, copy and characteristic inlet/outlet boundaries),
// written with the library's own host arithmetic so that the probe matches
// it; the time step is the user's own (its minimum is what min_dt must give).
//
//   dropin_phases <mesh name without .msh> <state file> <output directory>
//
// The state file holds 4 doubles per cell over all cells (the reference
// layout); each output is written raw as <output directory>/<name>.f64.
#define FVH_HD inline
#include "fvh_physics.h"

#include <fvhyper/explicit.h>
#include <fvhyper/mesh.h>
#include <fvhyper/parallel.h>
#include <fvhyper/post.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace fvhyper {

const int vars = 4;
const std::vector<std::string> var_names = {"rho", "rhou", "rhov", "rhoe"};

namespace solver {
const bool do_calc_gradients = true;
const bool do_calc_limiters = true;
const bool linear_interpolate = true;
const bool diffusive_gradients = false;
const bool global_dt = true;
const bool smooth_residuals = false;
}  // namespace solver

namespace {
const double kGamma = 1.4;
void bc_with(int kind, double* b, double* q, double* n) {
    fvh::BcParam<double> p{kind, 1.4, 3.0, 0.0, 1.0};
    fvh::boundary(p, kGamma, b, q, n[0], n[1]);
}
void bc_wall(double* b, double* q, double* n) { bc_with(fvh::BC_MIRROR, b, q, n); }
void bc_copy(double* b, double* q, double* n) { bc_with(fvh::BC_COPY, b, q, n); }
void bc_inout(double* b, double* q, double* n) { bc_with(fvh::BC_INLET_OUTLET, b, q, n); }
}  // namespace

void generate_initial_solution(std::vector<double>& v, const mesh& m) {
    for (size_t i = 0; i < m.cellsAreas.size(); ++i) {
        v[4 * i] = 1.4;
        v[4 * i + 1] = 1.4 * 3.0;
        v[4 * i + 2] = 0.0;
        v[4 * i + 3] = 1.0 / (kGamma - 1) + 0.5 * 1.4 * 9.0;
    }
}

double limiter_func(const double& r) { return fvh::phi<double, fvh::PHI_VENKAT>(r, 2.0); }

void calc_flux(double* f, const double* qi, const double* qj, const double*, const double*, const double* n) {
    fvh::roe_flux<double, true>(kGamma, f, qi, qj, n[0], n[1]);
}

// a local time step of the user's own: CFL 0.8 area over the sum of the
// edge spectral radii times lengths, the same value for every variable
void calc_dt(std::vector<double>& dt, const std::vector<double>& q, mesh& m) {
    std::vector<double> s(m.cellsAreas.size(), 0.0);
    for (uint e = 0; e < m.edgesLengths.size(); ++e) {
        const uint i = m.edgesCells(e, 0), j = m.edgesCells(e, 1);
        const double ei = fvh::eig_sum(kGamma, &q[4 * i], m.edgesNormalsX[e], m.edgesNormalsY[e]);
        const double ej = fvh::eig_sum(kGamma, &q[4 * j], m.edgesNormalsX[e], m.edgesNormalsY[e]);
        const double l = (ei > ej ? ei : ej) * m.edgesLengths[e];
        s[i] += l;
        s[j] += l;
    }
    for (size_t i = 0; i < s.size(); ++i)
        for (int k = 0; k < 4; ++k) dt[4 * i + k] = 0.8 * m.cellsAreas[i] / s[i];
}

namespace boundaries {
std::map<std::string, void (*)(double*, double*, double*)> bounds = {
    {"wall", bc_wall}, {"transmissive", bc_copy}, {"inlet", bc_inout}, {"outlet", bc_inout}};
}
namespace post {
std::map<std::string, void (*)(double*, double*)> extra_scalars = {};
std::map<std::string, void (*)(double*, double*)> extra_vectors = {};
}  // namespace post

}  // namespace fvhyper

using namespace fvhyper;

static void dump(const std::string& dir, const std::string& name, const std::vector<double>& v) {
    std::FILE* f = std::fopen((dir + "/" + name + ".f64").c_str(), "wb");
    if (!f || std::fwrite(v.data(), 8, v.size(), f) != v.size()) {
        std::cerr << "cannot write " << name << "\n";
        std::exit(2);
    }
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 4) {
        std::cerr << "usage: dropin_phases <mesh> <state file> <out dir>\n";
        return 2;
    }
    mpi_wrapper pool;
    mesh m;
    m.read_file(argv[1], pool);
    const size_t n = 4 * m.cellsAreas.size();
    std::vector<double> q(n);
    {
        std::ifstream in(argv[2], std::ios::binary);
        in.read((char*)q.data(), (std::streamsize)(n * 8));
        if (!in) { std::cerr << "short state file\n"; return 2; }
    }
    const std::string out = argv[3];
    std::vector<double> gx(n), gy(n), lim(n), qmin(n), qmax(n), qt(n), dt(n), qk(n);
    calc_gradients(gx, gy, q, m);
    calc_limiters(lim, qmin, qmax, q, gx, gy, m);
    std::vector<double> lim5(n);
    const mesh& cm = m;
    calc_limiters(lim5, q, gx, gy, cm);          // the header's form
    std::vector<double> gxc(n), gyc(n);
    calc_gradients(gxc, gyc, q, cm);            // the header's form
    calc_time_derivatives(qt, q, gx, gy, lim, m);
    calc_dt(dt, q, m);
    dump(out, "dt_local", dt);
    min_dt(dt, m);
    validate_dt(dt, pool);
    std::vector<double> q0 = q;
    update_cells(qk, q0, qt, dt, 0.506);
    dump(out, "qk_cells", qk);
    update_bounds(qk, gx, gy, lim, m);
    update_comms(qk, m);
    std::vector<double> R(4);
    calc_residuals(R.data(), qt, m, pool);
    std::vector<double> qt2(n), gx2(n), gy2(n), qmin2(n), qmax2(n), lim2(n);
    complete_calc_qt(qt2, q, gx2, gy2, qmin2, qmax2, lim2, m, pool);
    dump(out, "gx", gx);
    dump(out, "gy", gy);
    dump(out, "gx_const", gxc);
    dump(out, "gy_const", gyc);
    dump(out, "lim", lim);
    dump(out, "lim_header", lim5);
    dump(out, "qmin", qmin);
    dump(out, "qmax", qmax);
    dump(out, "qt", qt);
    dump(out, "dt", dt);
    dump(out, "qk", qk);
    dump(out, "R", R);
    dump(out, "qt_complete", qt2);
    dump(out, "lim_complete", lim2);
    std::cout << "phases ok " << m.nRealCells << " cells\n";
    return 0;
}
