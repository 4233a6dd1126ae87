// fvh_check.cpp -- host entry points over the device physics, for the CPU
// test suite only (libfvh_check.so): the edge flux, its hand-written reverse
// and its dual-number forward derivative; the boundary functions likewise.
// Nothing in the solver path calls these.
#include "fvh_adjoint.h"

using namespace fvh;

extern "C" {

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

double fvhc_phi_deriv(int kind, double y, double yt) {
    switch (kind) {
        case PHI_MICHALAK: return phi_deriv<double, PHI_MICHALAK>(y, yt);
        case PHI_VENKAT: return phi_deriv<double, PHI_VENKAT>(y, yt);
        case PHI_MINMOD: return phi_deriv<double, PHI_MINMOD>(y, yt);
        default: return 0.0;
    }
}

}  // extern "C"
