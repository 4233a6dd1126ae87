// fvh_cstep.cpp -- the explicit path of fvh_oracle.c restated as a C++
// template on the number type, for complex-step derivatives.  TEST
// INFRASTRUCTURE ONLY (libcstep.so): the product never links or calls it.
//
// Instantiated on double it must reproduce fvh_oracle.c's run() bit for bit
// (same operations, same order; a CPU test checks this), which ties it to
// the compiled reference.  Instantiated on `cs`, a complex number whose
// comparisons, min/max and |.| act on the real part (the usual complex-step
// conventions), a state q0 + i h v with h = 1e-30 gives the derivative of
// the objective along v as Im(J)/h, exact to rounding, on whatever branch
// the real evaluation takes.  That is the check of the GPU adjoint.
//
// Functions follow fvh_oracle.c, which cites the reference: calc_dt kinds
// (forwardStep/main.cpp:140-185, naca/main.cpp:142-183), min_dt
// (explicit.cpp:555-564), calc_gradients (explicit.cpp:112-171),
// calc_limiters (explicit.cpp:174-258), calc_time_derivatives
// (explicit.cpp:261-364), update_cells (explicit.cpp:368-378),
// update_bounds (explicit.cpp:380-417), run (explicit.cpp:657-817).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "fvh_oracle.h"

namespace {

constexpr int V = 4;

struct cs {
    double re, im;
    cs() : re(0.0), im(0.0) {}
    cs(double a) : re(a), im(0.0) {}
    cs(double a, double b) : re(a), im(b) {}
};
inline cs operator+(cs a, cs b) { return cs(a.re + b.re, a.im + b.im); }
inline cs operator-(cs a, cs b) { return cs(a.re - b.re, a.im - b.im); }
inline cs operator-(cs a) { return cs(-a.re, -a.im); }
inline cs operator*(cs a, cs b) { return cs(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
inline cs operator/(cs a, cs b) {
    const double d = b.re * b.re + b.im * b.im;
    return cs((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
inline cs& operator+=(cs& a, cs b) { a = a + b; return a; }
inline cs& operator-=(cs& a, cs b) { a = a - b; return a; }
inline cs& operator*=(cs& a, cs b) { a = a * b; return a; }
inline bool operator<(cs a, cs b) { return a.re < b.re; }
inline bool operator>(cs a, cs b) { return a.re > b.re; }
inline bool operator<=(cs a, cs b) { return a.re <= b.re; }
inline bool operator>=(cs a, cs b) { return a.re >= b.re; }
// principal square root, imaginary part without cancellation
inline cs sqrt(cs a) {
    if (a.im == 0.0 && a.re >= 0.0) return cs(::sqrt(a.re), 0.0);
    const double r = hypot(a.re, a.im);
    if (a.re >= 0.0) {
        const double s = ::sqrt(0.5 * (r + a.re));
        return cs(s, a.im / (2.0 * s));
    }
    const double t = ::sqrt(0.5 * (r - a.re));
    return cs(fabs(a.im) / (2.0 * t), a.im < 0.0 ? -t : t);
}
inline cs fabs(cs a) { return a.re < 0.0 ? -a : a; }
inline double re(double a) { return a; }
inline double re(cs a) { return a.re; }

template <class T> inline T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> inline T smax(T a, T b) { return (a < b) ? b : a; }
using ::fabs;
using ::sqrt;
// smoothed kinks (the oracle's kinks mode): |x|_w, min_w, max_w; analytic,
// so the complex step differentiates them
template <class T> inline T kabs(T x, double w) { return x * x / sqrt(x * x + T(w) * T(w)); }
template <class T> inline T kmin(T a, T b, double w) { return T(0.5) * (a + b - kabs(a - b, w)); }
template <class T> inline T kmax(T a, T b, double w) { return T(0.5) * (a + b + kabs(a - b, w)); }

template <class T>
inline T pressure(double g, const T* q) {
    return (T(g) - T(1.0)) * (q[3] - T(0.5) / q[0] * (q[1] * q[1] + q[2] * q[2]));
}

template <class T>
void flux(const orc_config* cf, T* f, const T* qi, const T* qj, const T* n) {
    const T g = T(cf->gamma);
    const T n0 = n[0], n1 = n[1];
    T pi = pressure(cf->gamma, qi);
    T pj = pressure(cf->gamma, qj);
    T Vi = (qi[1] * n0 + qi[2] * n1) / qi[0];
    T Vj = (qj[1] * n0 + qj[2] * n1) / qj[0];
    if (cf->flux == ORC_FLUX_ROE_JOINED) {
        f[0] = (qi[0] * Vi + qj[0] * Vj) * T(0.5);
        f[1] = (qi[1] * Vi + pi * n0 + qj[1] * Vj + pj * n0) * T(0.5);
        f[2] = (qi[2] * Vi + pi * n1 + qj[2] * Vj + pj * n1) * T(0.5);
        f[3] = ((qi[3] + pi) * Vi + (qj[3] + pj) * Vj) * T(0.5);
    } else {
        f[0] = qi[0] * Vi;
        f[1] = qi[1] * Vi + pi * n0;
        f[2] = qi[2] * Vi + pi * n1;
        f[3] = (qi[3] + pi) * Vi;
        f[0] += qj[0] * Vj;
        f[1] += qj[1] * Vj + pj * n0;
        f[2] += qj[2] * Vj + pj * n1;
        f[3] += (qj[3] + pj) * Vj;
        for (int i = 0; i < 4; ++i) f[i] *= T(0.5);
    }
    const T pL = pi, pR = pj;
    const T uL = qi[1] / qi[0];
    const T uR = qj[1] / qj[0];
    const T vL = qi[2] / qi[0];
    const T vR = qj[2] / qj[0];
    const T srhoL = sqrt(qi[0]);
    const T srhoR = sqrt(qj[0]);
    const T rho = srhoR * srhoL;
    const T u = (uL * srhoL + uR * srhoR) / (srhoL + srhoR);
    const T v = (vL * srhoL + vR * srhoR) / (srhoL + srhoR);
    const T h = ((qi[3] + pL) / qi[0] * srhoL + (qj[3] + pR) / qj[0] * srhoR) / (srhoL + srhoR);
    const T q2 = u * u + v * v;
    const T c = sqrt((g - T(1.)) * (h - T(0.5) * q2));
    const T Vn = u * n0 + v * n1;
    const T VR = uR * n0 + vR * n1;
    const T VL = uL * n0 + vL * n1;
    const double w = cf->kwv;
    const T lcm = cf->kinks ? kabs(kmin(Vn - c, VL - c, w), w) : fabs(smin(Vn - c, VL - c));
    const T lc = cf->kinks ? kabs(Vn, w) : fabs(Vn);
    const T lcp = cf->kinks ? kabs(kmax(Vn + c, VR + c, w), w) : fabs(smax(Vn + c, VR + c));
    const T kF1 = lcm * ((pR - pL) - rho * c * (VR - VL)) / (T(2.) * c * c);
    const T kF234_0 = lc * ((qj[0] - qi[0]) - (pR - pL) / (c * c));
    const T kF234_1 = lc * rho;
    const T kF5 = lcp * ((pR - pL) + rho * c * (VR - VL)) / (T(2.) * c * c);
    f[0] -= T(0.5) * (kF1 + kF234_0 + kF5);
    f[1] -= T(0.5) * (kF1 * (u - c * n0) + kF234_0 * u + kF234_1 * (uR - uL - (VR - VL) * n0) + kF5 * (u + c * n0));
    f[2] -= T(0.5) * (kF1 * (v - c * n1) + kF234_0 * v + kF234_1 * (vR - vL - (VR - VL) * n1) + kF5 * (v + c * n1));
    f[3] -= T(0.5) * (kF1 * (h - c * Vn) + kF234_0 * q2 * T(0.5) +
                      kF234_1 * (u * (uR - uL) + v * (vR - vL) - Vn * (VR - VL)) + kF5 * (h + c * Vn));
}

template <class T>
T phi(const orc_config* cf, T y) {
    switch (cf->phi) {
    case ORC_PHI_ZERO:
        return T(0.);
    case ORC_PHI_MICHALAK: {
        const T yt = T(cf->phi_yt);
        if (y >= yt) return T(1.0);
        const T a = T(1.0) / (yt * yt) - T(2.0) / (yt * yt * yt);
        const T b = T(-3.0 / 2.0) * a * yt - T(0.5) / yt;
        return a * y * y * y + b * y * y + y;
    }
    case ORC_PHI_VENKAT:
        return (y * y + T(2.0) * y) / (y * y + y + T(2.0));
    default:
        return smax(T(0.), smin(T(1.), y));
    }
}

template <class T>
void bc(const orc_config* cf, int slot, T* b, const T* q, const T* nd) {
    const int kind = cf->bc_kind[slot];
    const double gd = cf->gamma;
    const T g = T(gd);
    const T n0 = nd[0], n1 = nd[1];
    if (kind == ORC_BC_COPY) {
        b[0] = q[0]; b[1] = q[1]; b[2] = q[2]; b[3] = q[3];
        return;
    }
    if (kind == ORC_BC_MIRROR) {
        b[0] = q[0];
        b[1] = q[1] - T(2.0) * n0 * (n0 * q[1] + n1 * q[2]);
        b[2] = q[2] - T(2.0) * n1 * (n0 * q[1] + n1 * q[2]);
        b[3] = q[3];
        return;
    }
    const double* P = cf->bc_param[slot];
    T bv[4];
    bv[0] = T(P[0]);
    bv[1] = bv[0] * T(P[1]);
    bv[2] = bv[0] * T(P[2]);
    bv[3] = T(P[3]) / (g - T(1)) + T(0.5) / bv[0] * (bv[1] * bv[1] + bv[2] * bv[2]);
    T U0 = q[1] / q[0];
    T U1 = q[2] / q[0];
    T u_norm = sqrt(U0 * U0 + U1 * U1);
    T u_dot_n = U0 * n0 + U1 * n1;
    T p = pressure(gd, q);
    T c = sqrt(g * p / q[0]);
    T mach = u_norm / c;
    if (mach > T(1)) {
        if (u_dot_n < T(0)) {
            b[0] = bv[0]; b[1] = bv[1]; b[2] = bv[2]; b[3] = bv[3];
        } else {
            b[0] = q[0]; b[1] = q[1]; b[2] = q[2]; b[3] = q[3];
        }
        return;
    }
    T pa = pressure(gd, bv);
    T rhoa = bv[0];
    T ua = bv[1] / bv[0];
    T va = bv[2] / bv[0];
    T pd = p;
    T rhod = q[0];
    T ud = q[1] / q[0];
    T rho0 = q[0];
    T c0 = c;
    T vd = q[2] / q[0];
    if (u_dot_n < T(0)) {
        if (kind == ORC_BC_INLET_OUTLET) {
            T pb = pd;
            b[0] = bv[0]; b[1] = bv[1]; b[2] = bv[2];
            b[3] = pb / (g - T(1.)) + T(0.5) / b[0] * (b[1] * b[1] + b[2] * b[2]);
        } else {
            T pb = T(0.5) * (pa + pd - rho0 * c0 * (n0 * (ua - ud) + n1 * (va - vd)));
            b[0] = rhoa + (pb - pa) / (c0 * c0);
            b[1] = b[0] * (ua - n0 * (pa - pb) / (rho0 * c0));
            b[2] = b[1] * (va - n1 * (pa - pb) / (rho0 * c0));
            b[3] = pb / (g - T(1)) + T(0.5) / b[0] * (b[1] * b[1] + b[2] * b[2]);
        }
    } else {
        T pb = pa;
        b[0] = rhod + (pb - pd) / (c0 * c0);
        b[1] = b[0] * (ud + n0 * (pd - pb) / (rho0 * c0));
        b[2] = b[1] * (va + n1 * (pd - pb) / (rho0 * c0));
        b[3] = pb / (g - T(1.)) + T(0.5) / b[0] * (b[1] * b[1] + b[2] * b[2]);
    }
}

// Mesh geometry over the number type: the node coordinates may carry a
// complex perturbation.  Same operations and order as orc_mesh_build
// (fvh_oracle.c, after mesh.cpp:190-260 and the ghost cells of
// mesh.cpp:623-669), so on double it reproduces the oracle's arrays.
template <class T>
struct Geo {
    std::vector<T> ccx, ccy, area;            // all cells, ghosts included
    std::vector<T> ecx, ecy, enx, eny, elen;  // edges
};

template <class T>
void geometry(const orc_mesh* m, const T* X, const T* Y, Geo<T>& G) {
    const int32_t NC = m->n_cells, NR = m->n_real, E = m->n_edges;
    G.ccx.assign(NC, T(0.)); G.ccy.assign(NC, T(0.)); G.area.assign(NC, T(0.));
    G.ecx.assign(E, T(0.)); G.ecy.assign(E, T(0.)); G.enx.assign(E, T(0.)); G.eny.assign(E, T(0.));
    G.elen.assign(E, T(0.));
    for (int32_t c = 0; c < NR; ++c) {
        T cx = T(0.0), cy = T(0.0);
        const int n = m->cell_tri[c] ? 3 : 4;
        for (int j = 0; j < n; ++j) {
            cx += X[m->cell_nodes[4 * c + j]] / T((double)n);
            cy += Y[m->cell_nodes[4 * c + j]] / T((double)n);
        }
        G.ccx[c] = cx;
        G.ccy[c] = cy;
    }
    for (int32_t e = 0; e < E; ++e) {
        const int32_t n0 = m->edge_nodes[2 * e], n1 = m->edge_nodes[2 * e + 1];
        G.ecx[e] = (X[n1] + X[n0]) * T(0.5);
        G.ecy[e] = (Y[n1] + Y[n0]) * T(0.5);
        const T dex = X[n1] - X[n0];
        const T dey = Y[n1] - Y[n0];
        const T l = sqrt(dex * dex + dey * dey);
        G.elen[e] = l;
        G.enx[e] = -dey / l;
        G.eny[e] = dex / l;
        const int32_t c0 = m->edge_cells[2 * e];
        const T dot = G.enx[e] * (G.ecx[e] - G.ccx[c0]) + G.eny[e] * (G.ecy[e] - G.ccy[c0]);
        if (dot < T(0.)) {
            G.enx[e] *= T(-1.);
            G.eny[e] *= T(-1.);
        }
    }
    for (int32_t c = 0; c < NR; ++c) {
        const int32_t* cn = &m->cell_nodes[4 * c];
        const T x1 = X[cn[0]], x2 = X[cn[1]], x3 = X[cn[2]];
        const T y1 = Y[cn[0]], y2 = Y[cn[1]], y3 = Y[cn[2]];
        if (m->cell_tri[c]) {
            G.area[c] = T(0.5) * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2));
        } else {
            const T x4 = X[cn[3]], y4 = Y[cn[3]];
            G.area[c] = T(0.5) * fabs(x1 * (y2 - y3) + x2 * (y3 - y1) + x3 * (y1 - y2)) +
                        T(0.5) * fabs(x1 * (y3 - y4) + x3 * (y4 - y1) + x4 * (y1 - y3));
        }
    }
    for (int32_t b = 0; b < m->n_bnd; ++b) {
        const int32_t e = m->bnd_edge[b];
        const int32_t cell = m->edge_cells[2 * e], g = NR + b;
        const T dx = G.ecx[e] - G.ccx[cell];
        const T dy = G.ecy[e] - G.ccy[cell];
        const T dist = sqrt(dx * dx + dy * dy);
        G.area[g] = G.area[cell];
        G.ccx[g] = G.ecx[e] + dist * G.enx[e];
        G.ccy[g] = G.ecy[e] + dist * G.eny[e];
    }
}

template <class T>
void calc_dt(const orc_mesh* m, const Geo<T>& G, const orc_config* cf, const T* q, T* dt) {
    const int32_t N = m->n_cells;
    const double gd = cf->gamma;
    const T g = T(gd);
    if (cf->dt_kind == ORC_DT_FIXED) {
        for (int32_t i = 0; i < N; ++i) dt[i] = T(cf->dt_fixed);
        return;
    }
    if (cf->dt_kind == ORC_DT_SUM_CFL) {
        for (int32_t i = 0; i < N; ++i) dt[i] = T(0.);
        for (int32_t e = 0; e < m->n_edges; ++e) {
            int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
            T le = G.elen[e], n0 = G.enx[e], n1 = G.eny[e];
            const T* qi = &q[V * i];
            const T* qj = &q[V * j];
            T ci = sqrt(pressure(gd, qi) * g / qi[0]);
            T cj = sqrt(pressure(gd, qj) * g / qj[0]);
            T ei, ej, center;
            if (cf->kinks) {
                ei = ci + kabs(qi[1] / qi[0] * n0 + qi[2] / qi[0] * n1, cf->kwv);
                ej = cj + kabs(qj[1] / qj[0] * n0 + qj[2] / qj[0] * n1, cf->kwv);
                center = kmax(ei, ej, cf->kwv);
            } else {
                ei = ci + fabs(qi[1] / qi[0] * n0 + qi[2] / qi[0] * n1);
                ej = cj + fabs(qj[1] / qj[0] * n0 + qj[2] / qj[0] * n1);
                center = smax(ei, ej);
            }
            dt[i] += center * le;
            dt[j] += center * le;
        }
        for (int32_t i = 0; i < N; ++i) dt[i] = T(cf->cfl) * G.area[i] / dt[i];
        return;
    }
    for (int32_t i = 0; i < N; ++i) dt[i] = T(1.0);
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
        T le = G.elen[e], n0 = G.enx[e], n1 = G.eny[e];
        const T* qi = &q[V * i];
        const T* qj = &q[V * j];
        T ci = sqrt(pressure(gd, qi) * g / qi[0]);
        T cj = sqrt(pressure(gd, qj) * g / qj[0]);
        T ei = (ci + fabs(qi[1] / qi[0])) * n0 + (ci + fabs(qi[2] / qi[0])) * n1;
        T ej = (cj + fabs(qj[1] / qj[0])) * n0 + (cj + fabs(qj[2] / qj[0])) * n1;
        T dti = T(cf->cfl) * (G.area[i] / le) / fabs(ei);
        T dtj = T(cf->cfl) * (G.area[j] / le) / fabs(ej);
        dt[i] = smin(dt[i], dti);
        dt[j] = smin(dt[j], dtj);
    }
}

template <class T>
void min_dt(const orc_mesh* m, T* dt) {
    T mn = dt[0];
    for (int32_t i = 0; i < m->n_cells; ++i) mn = smin(mn, dt[i]);
    for (int32_t i = 0; i < m->n_cells; ++i) dt[i] = mn;
}

// the p-norm global step of the oracle's kinks mode: the real part as the
// oracle forms it (fixed-point sum), the imaginary part its derivative
// sum_i (dt / dt_i)^129 Im(dt_i)
inline double soft_value(const orc_mesh* m, const double* re_dt) {
    const int32_t N = m->n_cells;
    double mn = re_dt[0];
    for (int32_t i = 0; i < N; ++i) mn = smin(mn, re_dt[i]);
    unsigned __int128 S = 0;
    for (int32_t i = 0; i < N; ++i) {
        double y = mn / re_dt[i];
        if (!(y >= 0.0 && y <= 1.0)) { S += (unsigned __int128)1 << 100; continue; }
        for (int k = 0; k < 7; ++k) y = y * y;
        if (y >= 1.0) S += (unsigned __int128)1 << 64;
        else S += (unsigned __int128)(uint64_t)(y * 18446744073709551616.0);
    }
    double root = (double)(uint64_t)(S >> 64) + (double)(uint64_t)S * 0x1p-64;
    for (int k = 0; k < 7; ++k) root = ::sqrt(root);
    return (S >> 100) != 0 ? NAN : mn / root;
}
inline void min_dt_soft(const orc_mesh* m, double* dt) {
    const double v = soft_value(m, dt);
    for (int32_t i = 0; i < m->n_cells; ++i) dt[i] = v;
}
inline void min_dt_soft(const orc_mesh* m, cs* dt) {
    const int32_t N = m->n_cells;
    std::vector<double> r(N);
    for (int32_t i = 0; i < N; ++i) r[i] = dt[i].re;
    const double v = soft_value(m, r.data());
    double im = 0.0;
    for (int32_t i = 0; i < N; ++i) {
        double y = v / r[i], w = y;
        for (int k = 0; k < 7; ++k) w = w * w;
        im += w * y * dt[i].im;
    }
    for (int32_t i = 0; i < N; ++i) dt[i] = cs(v, im);
}

template <class T>
void gradients(const orc_mesh* m, const Geo<T>& G, const T* q, T* gx, T* gy) {
    const int64_t n = (int64_t)V * m->n_cells;
    for (int64_t i = 0; i < n; ++i) { gx[i] = T(0.); gy[i] = T(0.); }
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
        T nx = G.enx[e], ny = G.eny[e], le = G.elen[e];
        T dxif = G.ecx[e] - G.ccx[i];
        T dyif = G.ecy[e] - G.ccy[i];
        T dif = sqrt(dxif * dxif + dyif * dyif);
        T dxij = G.ccx[i] - G.ccx[j];
        T dyij = G.ccy[i] - G.ccy[j];
        T dij = sqrt(dxij * dxij + dyij * dyij);
        T gf = dif / dij;
        if (i != j) {
            T f[V];
            for (int k = 0; k < V; ++k) f[k] = (q[V * i + k] * (T(1.0) - gf) + q[V * j + k] * gf) * le;
            for (int k = 0; k < V; ++k) {
                gx[V * i + k] += f[k] * nx;
                gy[V * i + k] += f[k] * ny;
                gx[V * j + k] -= f[k] * nx;
                gy[V * j + k] -= f[k] * ny;
            }
        }
    }
    for (int32_t i = 0; i < m->n_real; ++i) {
        const T invA = T(1.) / G.area[i];
        for (int k = 0; k < V; ++k) { gx[V * i + k] *= invA; gy[V * i + k] *= invA; }
    }
    for (int32_t i = m->n_real; i < m->n_cells; ++i)
        for (int k = 0; k < V; ++k) { gx[V * i + k] = T(0.); gy[V * i + k] = T(0.); }
}

template <class T>
void limiters(const orc_mesh* m, const Geo<T>& G, const orc_config* cf, const T* q, const T* gx, const T* gy,
              T* lim, T* qmin, T* qmax) {
    const int64_t n = (int64_t)V * m->n_cells;
    for (int64_t i = 0; i < n; ++i) { lim[i] = T(1.); qmin[i] = q[i]; qmax[i] = q[i]; }
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
        for (int k = 0; k < V; ++k) {
            if (cf->kinks) {
                const double w = cf->kw[k];
                qmin[V * i + k] = kmin(qmin[V * i + k], q[V * j + k], w);
                qmin[V * j + k] = kmin(qmin[V * j + k], q[V * i + k], w);
                qmax[V * i + k] = kmax(qmax[V * i + k], q[V * j + k], w);
                qmax[V * j + k] = kmax(qmax[V * j + k], q[V * i + k], w);
                continue;
            }
            qmin[V * i + k] = smin(qmin[V * i + k], q[V * j + k]);
            qmin[V * j + k] = smin(qmin[V * j + k], q[V * i + k]);
            qmax[V * i + k] = smax(qmax[V * i + k], q[V * j + k]);
            qmax[V * j + k] = smax(qmax[V * j + k], q[V * i + k]);
        }
    }
    const T tol = T(1e-15);
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t ids[2] = {m->edge_cells[2 * e], m->edge_cells[2 * e + 1]};
        for (int s = 0; s < 2; ++s) {
            int32_t id = ids[s];
            if (id >= m->n_real) continue;
            T dx = G.ecx[e] - G.ccx[id];
            T dy = G.ecy[e] - G.ccy[id];
            T sqrt_area = sqrt(G.area[id]);
            for (int k = 0; k < V; ++k) {
                T dqg = gx[V * id + k] * dx + gy[V * id + k] * dy;
                T dmax = qmax[V * id + k] - q[V * id + k];
                T dmin = qmin[V * id + k] - q[V * id + k];
                const T Ka = T(cf->K) * sqrt_area;
                const T K3a = Ka * Ka * Ka;
                const T d2 = (dmax - dmin) * (dmax - dmin);
                T sig;
                if (d2 <= K3a) {
                    sig = T(1.);
                } else if (d2 <= T(2) * K3a && cf->kinks) {
                    T yy = (d2 / K3a - T(1.0));
                    T y2 = yy * yy;
                    sig = T(1.0) - y2 * y2 * yy *
                          (T(126.0) + yy * (T(-420.0) + yy * (T(540.0) + yy * (T(-315.0) + T(70.0) * yy))));
                } else if (d2 <= T(2) * K3a) {
                    T yy = (d2 / K3a - T(1.0));
                    sig = T(2.0) * yy * yy * yy - T(3.0) * yy * yy + T(1.0);
                } else {
                    sig = T(0.);
                }
                T l = T(1.0);
                if (sig < T(1.0) && cf->kinks) {
                    const double w = cf->kw[k];
                    const T t = dqg / sqrt(dqg * dqg + T(w) * T(w));
                    const T d = T(0.5) * (dmax + dmin) + T(0.5) * (dmax - dmin) * t;
                    l = (d * d + T(2.0) * d * dqg + T(w) * T(w)) / (d * d + d * dqg + T(2.0) * dqg * dqg + T(w) * T(w));
                } else if (sig < T(1.0)) {
                    if (dqg > tol) l = phi(cf, dmax / dqg);
                    else if (dqg < -tol) l = phi(cf, dmin / dqg);
                    else l = T(1.0);
                }
                l = sig + (T(1.0) - sig) * l;
                lim[V * id + k] = cf->kinks ? kmin(lim[V * id + k], l, cf->kwl) : smin(lim[V * id + k], l);
            }
        }
    }
}

template <class T>
bool physical(double g, const T* q) {
    return q[0] > T(0) && pressure(g, q) > T(0);
}

template <class T>
void time_derivatives(const orc_mesh* m, const Geo<T>& G, const orc_config* cf, const T* q, const T* gx,
                      const T* gy, const T* lim, T* qt) {
    const int64_t n = (int64_t)V * m->n_cells;
    for (int64_t i = 0; i < n; ++i) qt[i] = T(0.);
    for (int32_t e = 0; e < m->n_edges; ++e) {
        int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
        T le = G.elen[e];
        T nn[2] = {G.enx[e], G.eny[e]};
        T cx = G.ecx[e], cy = G.ecy[e];
        T di0 = cx - G.ccx[i], di1 = cy - G.ccy[i];
        T dj0 = cx - G.ccx[j], dj1 = cy - G.ccy[j];
        T qi[V], qj[V], f[V];
        if (cf->lin) {
            for (int k = 0; k < V; ++k) {
                int64_t ki = V * (int64_t)i + k, kj = V * (int64_t)j + k;
                qi[k] = q[ki] + (gx[ki] * di0 + gy[ki] * di1) * lim[ki];
                qj[k] = q[kj] + (gx[kj] * dj0 + gy[kj] * dj1) * lim[kj];
            }
        } else {
            for (int k = 0; k < V; ++k) { qi[k] = q[V * i + k]; qj[k] = q[V * j + k]; }
        }
        if (cf->lin && cf->pos) {
            if (!physical(cf->gamma, qi))
                for (int k = 0; k < V; ++k) qi[k] = q[V * i + k];
            if (!physical(cf->gamma, qj))
                for (int k = 0; k < V; ++k) qj[k] = q[V * j + k];
        }
        flux(cf, f, qi, qj, nn);
        for (int k = 0; k < V; ++k) {
            qt[V * i + k] -= f[k] * le / G.area[i];
            qt[V * j + k] += f[k] * le / G.area[j];
        }
    }
    for (int32_t i = m->n_real; i < m->n_cells; ++i)
        for (int k = 0; k < V; ++k) qt[V * i + k] = T(0.);
}

template <class T>
void update_bounds(const orc_mesh* m, const Geo<T>& G, const orc_config* cf, T* q, const T* gx, const T* gy,
                   const T* lim) {
    for (int32_t b = 0; b < m->n_bnd; ++b) {
        int32_t e = m->bnd_edge[b];
        T nn[2] = {G.enx[e], G.eny[e]};
        int32_t ii = m->edge_cells[2 * e], ib = m->edge_cells[2 * e + 1];
        T qint[V];
        if (cf->lin) {
            T d0 = G.ecx[e] - G.ccx[ii];
            T d1 = G.ecy[e] - G.ccy[ii];
            for (int k = 0; k < V; ++k) {
                int64_t kk = V * (int64_t)ii + k;
                qint[k] = q[kk] + (gx[kk] * d0 + gy[kk] * d1) * lim[kk];
            }
            if (cf->pos && !physical(cf->gamma, qint))
                for (int k = 0; k < V; ++k) qint[k] = q[V * ii + k];
        } else {
            for (int k = 0; k < V; ++k) qint[k] = q[V * ii + k];
        }
        bc(cf, m->bnd_bc[b], &q[V * ib], qint, nn);
    }
}

// smooth_residuals (explicit.cpp:32-66): two Jacobi sweeps, eps = 0.6
template <class T>
void smooth_residuals(const orc_mesh* m, T* qt) {
    const int64_t n = (int64_t)V * m->n_cells;
    const double epsilon = 0.6;
    std::vector<T> qt0(qt, qt + n), sm(n);
    for (int it = 0; it < 2; ++it) {
        for (int64_t i = 0; i < n; ++i) sm[i] = T(0.);
        for (int32_t e = 0; e < m->n_edges; ++e) {
            const int64_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
            for (int k = 0; k < V; ++k) {
                sm[V * i + k] -= qt[V * j + k] * epsilon;
                sm[V * j + k] -= qt[V * i + k] * epsilon;
            }
        }
        for (int32_t i = 0; i < m->n_real; ++i) {
            const double ne = m->cell_tri[i] ? 3 : 4;
            for (int k = 0; k < V; ++k) qt[V * i + k] = (qt0[V * i + k] - sm[V * i + k]) / (1. + ne * epsilon);
        }
    }
}

// run() for a fixed number of steps (explicit.cpp:657-817 without output)
template <class T>
void run(const orc_mesh* m, const Geo<T>& G, const orc_config* cf, T* q, int steps) {
    const int64_t N = m->n_cells, n = V * N;
    std::vector<T> qk(n), qt(n), gx(n), gy(n), lim(n), qmin(n), qmax(n), dt(N);
    update_bounds(m, G, cf, q, gx.data(), gy.data(), lim.data());
    for (int step = 0; step < steps; ++step) {
        calc_dt(m, G, cf, q, dt.data());
        if (cf->gdt) {
            if (cf->kinks && cf->dt_kind != ORC_DT_FIXED) min_dt_soft(m, dt.data());
            else min_dt(m, dt.data());
        }
        for (int64_t i = 0; i < n; ++i) qk[i] = q[i];
        for (int s = 0; s < cf->n_alpha; ++s) {
            if (cf->grad) gradients(m, G, qk.data(), gx.data(), gy.data());
            if (cf->lim) limiters(m, G, cf, qk.data(), gx.data(), gy.data(), lim.data(), qmin.data(), qmax.data());
            time_derivatives(m, G, cf, qk.data(), gx.data(), gy.data(), lim.data(), qt.data());
            if (cf->smooth) smooth_residuals(m, qt.data());
            const T a = T(cf->alpha[s]);
            for (int64_t c = 0; c < N; ++c)
                for (int k = 0; k < V; ++k) qk[V * c + k] = q[V * c + k] + qt[V * c + k] * dt[c] * a;
            update_bounds(m, G, cf, qk.data(), gx.data(), gy.data(), lim.data());
        }
        for (int64_t i = 0; i < n; ++i) q[i] = qk[i];
    }
}

// J = sum_b p(ghost_b) (wx_b n_x + wy_b n_y) len_b (the GPU objective's form)
template <class T>
T objective(const orc_mesh* m, const Geo<T>& G, const orc_config* cf, const T* q, const double* wx,
            const double* wy) {
    T J = T(0.);
    for (int32_t b = 0; b < m->n_bnd; ++b) {
        const int32_t e = m->bnd_edge[b];
        const int32_t ghost = m->n_real + b;
        const T proj = T(wx ? wx[b] : 0.0) * G.elen[e] * G.enx[e] + T(wy ? wy[b] : 0.0) * G.elen[e] * G.eny[e];
        J += pressure(cf->gamma, &q[V * (int64_t)ghost]) * proj;
    }
    return J;
}

}  // namespace

extern "C" {

// double instantiation of the geometry: 5 arrays over all cells (cx, cy,
// area) then 5 over edges (cx, cy, nx, ny, len), packed in that order
int cst_geometry_f64(const orc_mesh* m, double* out) {
    Geo<double> G;
    geometry<double>(m, m->x, m->y, G);
    const int64_t NC = m->n_cells, E = m->n_edges;
    const std::vector<double>* cells[3] = {&G.ccx, &G.ccy, &G.area};
    const std::vector<double>* edges[5] = {&G.ecx, &G.ecy, &G.enx, &G.eny, &G.elen};
    for (int k = 0; k < 3; ++k) memcpy(out + k * NC, cells[k]->data(), sizeof(double) * NC);
    for (int k = 0; k < 5; ++k) memcpy(out + 3 * NC + k * E, edges[k]->data(), sizeof(double) * E);
    return 0;
}

// complex-step derivative of the geometry along node displacements (dx, dy):
// per edge (n_x, n_y, len, Green-Gauss weight as calc_gradients forms it,
// centre x, centre y), then per cell, ghosts included, (area, centre x, y)
int cst_geometry_jvp(const orc_mesh* m, const double* dx, const double* dy, double h, double* out) {
    std::vector<cs> X(m->n_nodes), Y(m->n_nodes);
    for (int32_t k = 0; k < m->n_nodes; ++k) {
        X[k] = cs(m->x[k], h * dx[k]);
        Y[k] = cs(m->y[k], h * dy[k]);
    }
    Geo<cs> G;
    geometry<cs>(m, X.data(), Y.data(), G);
    const int64_t E = m->n_edges;
    for (int32_t e = 0; e < m->n_edges; ++e) {
        const int32_t i = m->edge_cells[2 * e], j = m->edge_cells[2 * e + 1];
        const cs dxif = G.ecx[e] - G.ccx[i], dyif = G.ecy[e] - G.ccy[i];
        const cs dxij = G.ccx[i] - G.ccx[j], dyij = G.ccy[i] - G.ccy[j];
        const cs w = sqrt(dxif * dxif + dyif * dyif) / sqrt(dxij * dxij + dyij * dyij);
        const cs v[6] = {G.enx[e], G.eny[e], G.elen[e], w, G.ecx[e], G.ecy[e]};
        for (int k = 0; k < 6; ++k) out[6 * e + k] = v[k].im / h;
    }
    for (int32_t c = 0; c < m->n_cells; ++c) {
        out[6 * E + 3 * c] = G.area[c].im / h;
        out[6 * E + 3 * c + 1] = G.ccx[c].im / h;
        out[6 * E + 3 * c + 2] = G.ccy[c].im / h;
    }
    return 0;
}

// double instantiation: final state after `steps` steps (all cells) and J
int cst_run_f64(const orc_mesh* m, const orc_config* cf, const double* q0, int steps, const double* wx,
                const double* wy, double* q_out, double* J) {
    const int64_t n = (int64_t)V * m->n_cells;
    Geo<double> G;
    geometry<double>(m, m->x, m->y, G);
    std::vector<double> q(q0, q0 + n);
    run<double>(m, G, cf, q.data(), steps);
    if (q_out) memcpy(q_out, q.data(), sizeof(double) * n);
    if (J) *J = objective<double>(m, G, cf, q.data(), wx, wy);
    return 0;
}

// complex step along (dir in the initial state, dx/dy in the node
// coordinates; either may be NULL): J(q0 + i h dir; x + i h dx, y + i h dy)
// -> J (real part) and dJ = Im / h
int cst_derivative2(const orc_mesh* m, const orc_config* cf, const double* q0, const double* dir,
                    const double* dx, const double* dy, double h, int steps, const double* wx, const double* wy,
                    double* J, double* dJ) {
    const int64_t n = (int64_t)V * m->n_cells;
    std::vector<cs> X(m->n_nodes), Y(m->n_nodes);
    for (int32_t k = 0; k < m->n_nodes; ++k) {
        X[k] = cs(m->x[k], dx ? h * dx[k] : 0.0);
        Y[k] = cs(m->y[k], dy ? h * dy[k] : 0.0);
    }
    Geo<cs> G;
    geometry<cs>(m, X.data(), Y.data(), G);
    std::vector<cs> q(n);
    for (int64_t i = 0; i < n; ++i) q[i] = cs(q0[i], dir ? h * dir[i] : 0.0);
    run<cs>(m, G, cf, q.data(), steps);
    const cs j = objective<cs>(m, G, cf, q.data(), wx, wy);
    *J = j.re;
    *dJ = j.im / h;
    return 0;
}

int cst_derivative(const orc_mesh* m, const orc_config* cf, const double* q0, const double* dir, double h,
                   int steps, const double* wx, const double* wy, double* J, double* dJ) {
    return cst_derivative2(m, cf, q0, dir, nullptr, nullptr, h, steps, wx, wy, J, dJ);
}

}  // extern "C"
