// fvh_check.cpp -- host entry points over the device physics, for the CPU
// test suite only (libfvh_check.so): the edge flux, its hand-written reverse
// and its dual-number forward derivative; the boundary functions likewise.
// Nothing in the solver path calls these.
#include "fvh_adjoint.h"
#include "fvh_internal.h"

#include <string>
#include <vector>

using namespace fvh;

extern "C" {

// Reverse of the mesh geometry (mesh_geometry_vjp) on a mesh built from
// file-order arrays (as fvh_mesh_from_arrays takes them).  bars: per edge
// (n_x, n_y, len, Green-Gauss weight, centre x, centre y), then per cell,
// ghosts included, (area, centre x, centre y); gxy: 2 per node.
int fvhc_geometry_vjp(int n_nodes, const double* x, const double* y, int n_cells, const int* cells, int n_bnd,
                      const int* bnd, const int* tags, const double* bars, double* gxy) {
    int nt = 0;
    for (int b = 0; b < n_bnd; ++b) nt = tags[b] + 1 > nt ? tags[b] + 1 : nt;
    std::vector<std::string> names;
    for (int k = 0; k < nt; ++k) names.push_back("b" + std::to_string(k));
    fvh_mesh m;
    const int rc = build_mesh(m, n_nodes, x, y, n_cells, cells, n_bnd, bnd, tags, names);
    if (rc) return rc;
    const int E = m.n_edges(), N = m.n_cells;
    GeoBar b;
    b.resize(E, N);
    std::vector<double>* ev[6] = {&b.e_nx, &b.e_ny, &b.e_len, &b.e_w, &b.e_cx, &b.e_cy};
    std::vector<double>* cv[3] = {&b.c_area, &b.c_cx, &b.c_cy};
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < 6; ++k) (*ev[k])[e] = bars[6 * e + k];
    for (int c = 0; c < N; ++c)
        for (int k = 0; k < 3; ++k) (*cv[k])[c] = bars[6 * E + 3 * c + k];
    for (int i = 0; i < 2 * n_nodes; ++i) gxy[i] = 0.0;
    mesh_geometry_vjp(m, b, gxy);
    return 0;
}

void fvhc_roe(double g, int joined, const double* qi, const double* qj, const double* n, double* f) {
    if (joined) roe_flux<double, true>(g, f, qi, qj, n[0], n[1]);
    else roe_flux<double, false>(g, f, qi, qj, n[0], n[1]);
}

void fvhc_roe_vjp(double g, const double* qi, const double* qj, const double* n, const double* fb, double* qib,
                  double* qjb, double* nb) {
    roe_flux_vjp<double>(g, qi, qj, n[0], n[1], fb, qib, qjb, nb);
}

// df = J (dqi, dqj, dn)
void fvhc_roe_jvp(double g, int joined, const double* qi, const double* qj, const double* n, const double* dqi,
                  const double* dqj, const double* dn, double* df) {
    Dual<double> a[4], b[4], f[4];
    for (int k = 0; k < 4; ++k) { a[k] = Dual<double>(qi[k], dqi[k]); b[k] = Dual<double>(qj[k], dqj[k]); }
    const Dual<double> n0(n[0], dn[0]), n1(n[1], dn[1]);
    if (joined) roe_flux<Dual<double>, true>(Dual<double>(g), f, a, b, n0, n1);
    else roe_flux<Dual<double>, false>(Dual<double>(g), f, a, b, n0, n1);
    for (int k = 0; k < 4; ++k) df[k] = f[k].d;
}

static BcParam<double> bcp(int kind, const double* par) {
    BcParam<double> P;
    P.kind = kind; P.rho = par[0]; P.u = par[1]; P.v = par[2]; P.p = par[3];
    return P;
}

void fvhc_bc_vjp(int kind, const double* par, double g, const double* q, const double* n, const double* bb,
                 double* qb) {
    boundary_vjp<double>(bcp(kind, par), g, q, n[0], n[1], bb, qb);
}

void fvhc_bc(int kind, const double* par, double g, const double* q, const double* n, double* b) {
    boundary<double>(bcp(kind, par), g, b, q, n[0], n[1]);
}

// reverse and dual-number forward derivative of the time-step wave speeds
void fvhc_eig_vjp(int kind, double g, const double* q, const double* n, double eb, double* qb) {
    if (kind == 1) eig_sum_vjp<double>(g, q, n[0], n[1], eb, qb);
    else eig_edge_vjp<double>(g, q, n[0], n[1], eb, qb);
}

double fvhc_eig_jvp(int kind, double g, const double* q, const double* n, const double* dq) {
    Dual<double> a[4];
    for (int k = 0; k < 4; ++k) a[k] = Dual<double>(q[k], dq[k]);
    const Dual<double> e = kind == 1 ? eig_sum(Dual<double>(g), a, Dual<double>(n[0]), Dual<double>(n[1]))
                                     : eig_edge(Dual<double>(g), a, Dual<double>(n[0]), Dual<double>(n[1]));
    return e.d;
}

// adjoints of the edge normal through a boundary function and through the
// time-step wave speeds
void fvhc_bc_vjp_n(int kind, const double* par, double g, const double* q, const double* n, const double* bb,
                   double* nb) {
    boundary_vjp_n<double>(bcp(kind, par), g, q, n[0], n[1], bb, nb);
}

void fvhc_eig_vjp_n(int kind, double g, const double* q, const double* n, double eb, double* nb) {
    if (kind == 1) eig_sum_vjp_n<double>(q, n[0], n[1], eb, nb);
    else eig_edge_vjp_n<double>(g, q, eb, nb);
}

double fvhc_phi_deriv(int kind, double y, double yt) {
    switch (kind) {
        case PHI_MICHALAK: return phi_deriv<double, PHI_MICHALAK>(y, yt);
        case PHI_VENKAT: return phi_deriv<double, PHI_VENKAT>(y, yt);
        case PHI_MINMOD: return phi_deriv<double, PHI_MINMOD>(y, yt);
        default: return 0.0;
    }
}

}  // extern "C"
