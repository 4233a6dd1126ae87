// fvh_adjoint.h -- reverse-mode (vector-Jacobian) forms of the per-edge,
// per-boundary and per-cell functions of the explicit path.
//
// Each *_vjp function takes the adjoint of a function's outputs and adds the
// adjoint of its inputs (x_bar += J^T y_bar).  Derivatives are those of the
// branch the forward evaluation takes: the min/max of the entropy fix and of
// the limiter, |.|, the limiter's sigma ranges and the boundary regimes are
// differentiated piecewise, exactly as a complex-step or dual-number
// evaluation of the same code differentiates them.  The SM forms are the
// derivatives of the smoothed kinks (FVH_FLAG_SMOOTH_KINKS: |.|_w, min_w,
// max_w of fvh_physics.h).
#pragma once

#include "fvh_dual.h"
#include "fvh_physics.h"

namespace fvh {

// p = (g-1)(q3 - 0.5/q0 (q1^2 + q2^2))
template <class T>
FVH_HD void pressure_vjp(T g, const T* q, T pb, T* qb) {
    const T gm = g - T(1);
    const T r = T(1) / q[0];
    const T u = q[1] * r, v = q[2] * r;
    const T pg = pb * gm;
    qb[0] = qb[0] + pg * T(0.5) * (u * u + v * v);
    qb[1] = qb[1] - pg * u;
    qb[2] = qb[2] - pg * v;
    qb[3] = qb[3] + pg;
}

// Reverse of roe_flux (both spellings have the same derivative): given fb
// (adjoint of f[4]) add to qib, qjb (4 each) and nb[2].  SM: the smoothed
// entropy cure of width w (roe_flux<T, J, true>).
template <class T, bool SM = false>
FVH_HD void roe_flux_vjp(T g, const T* qi, const T* qj, T n0, T n1, const T* fb, T* qib, T* qjb, T* nb,
                         T w = T(0)) {
    // ---- forward recomputation (the values roe_flux forms) ----
    // The derivative is not required to round like roe_flux, so repeated
    // divisions become multiplications by a reciprocal.  The sign tests
    // below can only change where the rounding straddles a tie, where the
    // flux has no derivative anyway.
    const T r_i = T(1) / qi[0], r_j = T(1) / qj[0];
    const T pi = pressure(g, qi);
    const T pj = pressure(g, qj);
    const T uL = qi[1] * r_i, uR = qj[1] * r_j;
    const T vL = qi[2] * r_i, vR = qj[2] * r_j;
    const T Vi = uL * n0 + vL * n1;
    const T Vj = uR * n0 + vR * n1;
    const T sL = sqrt(qi[0]), sR = sqrt(qj[0]);
    const T rho = sR * sL;
    const T den = sL + sR;
    const T r_d = T(1) / den;
    const T u = (uL * sL + uR * sR) * r_d;
    const T v = (vL * sL + vR * sR) * r_d;
    const T hL = (qi[3] + pi) * r_i, hR = (qj[3] + pj) * r_j;
    const T h = (hL * sL + hR * sR) * r_d;
    const T q2 = u * u + v * v;
    const T c = sqrt((g - T(1)) * (h - T(0.5) * q2));
    const T V = u * n0 + v * n1;
    const T VR = Vj;
    const T VL = Vi;
    const T am = V - c, bm = VL - c;
    const bool m_is_b = bm < am;
    const T mm = SM ? kmin(am, bm, w) : (m_is_b ? bm : am);
    const T lcm = SM ? kabs(mm, w) : sabs(mm);
    const T lc = SM ? kabs(V, w) : sabs(V);
    const T ap = V + c, bp = VR + c;
    const bool p_is_b = ap < bp;
    const T MM = SM ? kmax(ap, bp, w) : (p_is_b ? bp : ap);
    const T lcp = SM ? kabs(MM, w) : sabs(MM);
    const T dp = pj - pi;
    const T dV = VR - VL;
    const T rc = rho * c;
    const T r_cc = T(1) / (c * c);
    const T r_c2 = T(0.5) * r_cc;
    const T c2x2 = T(2) * c * c;
    const T A1 = dp - rc * dV, A5 = dp + rc * dV;
    const T kF1 = lcm * A1 * r_c2;
    const T B0 = (qj[0] - qi[0]) - dp * r_cc;
    const T kF0 = lc * B0;
    const T kFr = lc * rho;
    const T kF5 = lcp * A5 * r_c2;
    // ---- reverse ----
    // f = F - 0.5 D, F = 0.5 (Fi + Fj)
    T Db[4], Fib[4];
    for (int k = 0; k < 4; ++k) { Db[k] = T(-0.5) * fb[k]; Fib[k] = T(0.5) * fb[k]; }
    // D0 = kF1 + kF0 + kF5
    // D1 = kF1 (u - c n0) + kF0 u + kFr (uR - uL - dV n0) + kF5 (u + c n0)
    // D2 = kF1 (v - c n1) + kF0 v + kFr (vR - vL - dV n1) + kF5 (v + c n1)
    // D3 = kF1 (h - c V) + kF0 q2/2 + kFr (u (uR-uL) + v (vR-vL) - V dV) + kF5 (h + c V)
    const T kF1b = Db[0] + Db[1] * (u - c * n0) + Db[2] * (v - c * n1) + Db[3] * (h - c * V);
    const T kF0b = Db[0] + Db[1] * u + Db[2] * v + Db[3] * q2 * T(0.5);
    const T kFrb = Db[1] * (uR - uL - dV * n0) + Db[2] * (vR - vL - dV * n1) +
                   Db[3] * (u * (uR - uL) + v * (vR - vL) - V * dV);
    const T kF5b = Db[0] + Db[1] * (u + c * n0) + Db[2] * (v + c * n1) + Db[3] * (h + c * V);
    const T ksum = kF1 + kF0 + kF5;
    T ub = Db[1] * ksum + Db[3] * kFr * (uR - uL);
    T vb = Db[2] * ksum + Db[3] * kFr * (vR - vL);
    T cb = (kF5 - kF1) * (Db[1] * n0 + Db[2] * n1 + Db[3] * V);
    T n0b = Db[1] * (c * (kF5 - kF1) - kFr * dV);
    T n1b = Db[2] * (c * (kF5 - kF1) - kFr * dV);
    T hb = Db[3] * (kF1 + kF5);
    T q2b = Db[3] * kF0 * T(0.5);
    T Vb = Db[3] * (c * (kF5 - kF1) - kFr * dV);
    T dVb = -kFr * (Db[1] * n0 + Db[2] * n1 + Db[3] * V);
    T uRb = kFr * (Db[1] + Db[3] * u);
    T uLb = -uRb;
    T vRb = kFr * (Db[2] + Db[3] * v);
    T vLb = -vRb;
    // kF1 = lcm A1 / c2x2 ; kF5 = lcp A5 / c2x2
    T lcmb = kF1b * A1 * r_c2;
    const T A1b = kF1b * lcm * r_c2;
    T lcpb = kF5b * A5 * r_c2;
    const T A5b = kF5b * lcp * r_c2;
    T c2x2b = -(kF1b * kF1 + kF5b * kF5) * r_c2;
    T dpb = A1b + A5b;
    T rcb = (A5b - A1b) * dV;
    dVb = dVb + (A5b - A1b) * rc;
    // kF0 = lc B0 ; kFr = lc rho
    T lcb = kF0b * B0 + kFrb * rho;
    const T B0b = kF0b * lc;
    T rhob = kFrb * lc;
    T qi0b = -B0b, qj0b = B0b;
    dpb = dpb - B0b * r_cc;
    cb = cb + B0b * T(2) * dp * r_cc / c;
    cb = cb + c2x2b * T(4) * c;
    rhob = rhob + rcb * c;
    cb = cb + rcb * rho;
    T pjb = dpb, pib = -dpb;
    T VRb = dVb, VLb = -dVb;
    // lcm = |min(V - c, VL - c)|, lc = |V|, lcp = |max(V + c, VR + c)|
    if (SM) {
        const T mmb = lcmb * kabs_d(mm, w);
        const T dm = kabs_d(am - bm, w);
        Vb = Vb + mmb * T(0.5) * (T(1) - dm);
        VLb = VLb + mmb * T(0.5) * (T(1) + dm);
        cb = cb - mmb;
        Vb = Vb + lcb * kabs_d(V, w);
        const T MMb = lcpb * kabs_d(MM, w);
        const T dp2 = kabs_d(ap - bp, w);
        Vb = Vb + MMb * T(0.5) * (T(1) + dp2);
        VRb = VRb + MMb * T(0.5) * (T(1) - dp2);
        cb = cb + MMb;
    } else {
        const T mmb = (mm < T(0)) ? -lcmb : lcmb;
        if (m_is_b) { VLb = VLb + mmb; cb = cb - mmb; } else { Vb = Vb + mmb; cb = cb - mmb; }
        Vb = Vb + ((V < T(0)) ? -lcb : lcb);
        const T MMb = (MM < T(0)) ? -lcpb : lcpb;
        if (p_is_b) { VRb = VRb + MMb; cb = cb + MMb; } else { Vb = Vb + MMb; cb = cb + MMb; }
    }
    // V, VR, VL
    ub = ub + Vb * n0; vb = vb + Vb * n1; n0b = n0b + Vb * u; n1b = n1b + Vb * v;
    uRb = uRb + VRb * n0; vRb = vRb + VRb * n1; n0b = n0b + VRb * uR; n1b = n1b + VRb * vR;
    uLb = uLb + VLb * n0; vLb = vLb + VLb * n1; n0b = n0b + VLb * uL; n1b = n1b + VLb * vL;
    // c = sqrt((g-1)(h - q2/2))
    const T cin = (g - T(1)) * cb / (T(2) * c);
    hb = hb + cin;
    q2b = q2b - cin * T(0.5);
    ub = ub + q2b * T(2) * u;
    vb = vb + q2b * T(2) * v;
    // h, u, v: Roe averages
    T sLb = (hb * (hL - h) + ub * (uL - u) + vb * (vL - v)) * r_d + rhob * sR;
    T sRb = (hb * (hR - h) + ub * (uR - u) + vb * (vR - v)) * r_d + rhob * sL;
    const T wL = sL * r_d, wR = sR * r_d;
    T hLb = hb * wL, hRb = hb * wR;
    uLb = uLb + ub * wL; uRb = uRb + ub * wR;
    vLb = vLb + vb * wL; vRb = vRb + vb * wR;
    qi0b = qi0b + sLb / (T(2) * sL);
    qj0b = qj0b + sRb / (T(2) * sR);
    // hL = (qi3 + pi)/qi0
    T qi3b = hLb * r_i, qj3b = hRb * r_j;
    pib = pib + hLb * r_i;
    pjb = pjb + hRb * r_j;
    qi0b = qi0b - hLb * hL * r_i;
    qj0b = qj0b - hRb * hR * r_j;
    T qi1b = uLb * r_i, qi2b = vLb * r_i;
    T qj1b = uRb * r_j, qj2b = vRb * r_j;
    qi0b = qi0b - (uLb * uL + vLb * vL) * r_i;
    qj0b = qj0b - (uRb * uR + vRb * vR) * r_j;
    // central flux, side i and side j
    T Vib = Fib[0] * qi[0] + Fib[1] * qi[1] + Fib[2] * qi[2] + Fib[3] * (qi[3] + pi);
    T Vjb = Fib[0] * qj[0] + Fib[1] * qj[1] + Fib[2] * qj[2] + Fib[3] * (qj[3] + pj);
    qi0b = qi0b + Fib[0] * Vi; qj0b = qj0b + Fib[0] * Vj;
    qi1b = qi1b + Fib[1] * Vi; qj1b = qj1b + Fib[1] * Vj;
    qi2b = qi2b + Fib[2] * Vi; qj2b = qj2b + Fib[2] * Vj;
    qi3b = qi3b + Fib[3] * Vi; qj3b = qj3b + Fib[3] * Vj;
    pib = pib + Fib[1] * n0 + Fib[2] * n1 + Fib[3] * Vi;
    pjb = pjb + Fib[1] * n0 + Fib[2] * n1 + Fib[3] * Vj;
    n0b = n0b + Fib[1] * (pi + pj);
    n1b = n1b + Fib[2] * (pi + pj);
    // Vi = (qi1 n0 + qi2 n1)/qi0
    qi1b = qi1b + Vib * n0 * r_i; qi2b = qi2b + Vib * n1 * r_i; qi0b = qi0b - Vib * Vi * r_i;
    n0b = n0b + Vib * uL + Vjb * uR;
    n1b = n1b + Vib * vL + Vjb * vR;
    qj1b = qj1b + Vjb * n0 * r_j; qj2b = qj2b + Vjb * n1 * r_j; qj0b = qj0b - Vjb * Vj * r_j;
    qib[0] = qib[0] + qi0b; qib[1] = qib[1] + qi1b; qib[2] = qib[2] + qi2b; qib[3] = qib[3] + qi3b;
    qjb[0] = qjb[0] + qj0b; qjb[1] = qjb[1] + qj1b; qjb[2] = qjb[2] + qj2b; qjb[3] = qjb[3] + qj3b;
    pressure_vjp(g, qi, pib, qib);
    pressure_vjp(g, qj, pjb, qjb);
    nb[0] = nb[0] + n0b;
    nb[1] = nb[1] + n1b;
}

// Reverse of eig_sum: e = c + |u.n|, c = sqrt(p g / rho); SM: of eig_sum_k
// (|u.n|_w).
template <class T, bool SM = false>
FVH_HD void eig_sum_vjp(T g, const T* q, T n0, T n1, T eb, T* qb, T w = T(0)) {
    const T p = pressure(g, q);
    const T X = p * g / q[0];
    const T c = sqrt(X);
    const T un = q[1] / q[0] * n0 + q[2] / q[0] * n1;
    const T Xb = eb / (T(2) * c);
    const T unb = SM ? eb * kabs_d(un, w) : ((un < T(0)) ? -eb : eb);
    pressure_vjp(g, q, Xb * g / q[0], qb);
    qb[0] = qb[0] - Xb * X / q[0] - unb * un / q[0];
    qb[1] = qb[1] + unb * n0 / q[0];
    qb[2] = qb[2] + unb * n1 / q[0];
}

// Reverse of eig_edge: e = (c + |u|) n0 + (c + |v|) n1.
template <class T>
FVH_HD void eig_edge_vjp(T g, const T* q, T n0, T n1, T eb, T* qb) {
    const T p = pressure(g, q);
    const T X = p * g / q[0];
    const T c = sqrt(X);
    const T u = q[1] / q[0], v = q[2] / q[0];
    const T Xb = eb * (n0 + n1) / (T(2) * c);
    const T ub = (u < T(0)) ? -eb * n0 : eb * n0;
    const T vb = (v < T(0)) ? -eb * n1 : eb * n1;
    pressure_vjp(g, q, Xb * g / q[0], qb);
    qb[0] = qb[0] - Xb * X / q[0] - (ub * u + vb * v) / q[0];
    qb[1] = qb[1] + ub / q[0];
    qb[2] = qb[2] + vb / q[0];
}

// Reverse of boundary(): b = BC(q, n); add J_q^T bb to qb (via dual numbers,
// one pass per input component; boundaries are few, so cost does not matter).
template <class T>
FVH_HD void boundary_vjp(const BcParam<T>& P, T g, const T* q, T n0, T n1, const T* bb, T* qb) {
    BcParam<Dual<T>> Pd;
    Pd.kind = P.kind;
    Pd.rho = Dual<T>(P.rho); Pd.u = Dual<T>(P.u); Pd.v = Dual<T>(P.v); Pd.p = Dual<T>(P.p);
    for (int j = 0; j < 4; ++j) {
        Dual<T> qd[4], bd[4];
        for (int k = 0; k < 4; ++k) qd[k] = Dual<T>(q[k], k == j ? T(1) : T(0));
        boundary(Pd, Dual<T>(g), bd, qd, Dual<T>(n0), Dual<T>(n1));
        T s = T(0);
        for (int k = 0; k < 4; ++k) s = s + bd[k].d * bb[k];
        qb[j] = qb[j] + s;
    }
}

// One column of the boundary function's Jacobian against bb: sum_k
// (d b_k / d x_j) bb_k with x = (q_0..q_3, n_0, n_1), by one dual-number
// evaluation.
template <class T>
FVH_HD T boundary_dir(const BcParam<T>& P, T g, const T* q, T n0, T n1, int j, const T* bb) {
    BcParam<Dual<T>> Pd;
    Pd.kind = P.kind;
    Pd.rho = Dual<T>(P.rho); Pd.u = Dual<T>(P.u); Pd.v = Dual<T>(P.v); Pd.p = Dual<T>(P.p);
    Dual<T> qd[4], bd[4];
    for (int k = 0; k < 4; ++k) qd[k] = Dual<T>(q[k], k == j ? T(1) : T(0));
    boundary(Pd, Dual<T>(g), bd, qd, Dual<T>(n0, j == 4 ? T(1) : T(0)), Dual<T>(n1, j == 5 ? T(1) : T(0)));
    T s = T(0);
    for (int k = 0; k < 4; ++k) s = s + bd[k].d * bb[k];
    return s;
}

// Adjoint of the edge normal through a boundary function: nb += (db/dn)^T bb
// (two dual-number evaluations, one per normal component).
template <class T>
FVH_HD void boundary_vjp_n(const BcParam<T>& P, T g, const T* q, T n0, T n1, const T* bb, T* nb) {
    BcParam<Dual<T>> Pd;
    Pd.kind = P.kind;
    Pd.rho = Dual<T>(P.rho); Pd.u = Dual<T>(P.u); Pd.v = Dual<T>(P.v); Pd.p = Dual<T>(P.p);
    for (int j = 0; j < 2; ++j) {
        Dual<T> qd[4], bd[4];
        for (int k = 0; k < 4; ++k) qd[k] = Dual<T>(q[k]);
        boundary(Pd, Dual<T>(g), bd, qd, Dual<T>(n0, j == 0 ? T(1) : T(0)), Dual<T>(n1, j == 1 ? T(1) : T(0)));
        T s = T(0);
        for (int k = 0; k < 4; ++k) s = s + bd[k].d * bb[k];
        nb[j] = nb[j] + s;
    }
}

// Normal part of the reverse of eig_sum (e = c + |u.n|): nb += eb d e / d n;
// SM: of eig_sum_k.
template <class T, bool SM = false>
FVH_HD void eig_sum_vjp_n(const T* q, T n0, T n1, T eb, T* nb, T w = T(0)) {
    const T un = q[1] / q[0] * n0 + q[2] / q[0] * n1;
    const T unb = SM ? eb * kabs_d(un, w) : ((un < T(0)) ? -eb : eb);
    nb[0] = nb[0] + unb * (q[1] / q[0]);
    nb[1] = nb[1] + unb * (q[2] / q[0]);
}

// Normal part of the reverse of eig_edge (e = (c+|u|) n0 + (c+|v|) n1).
template <class T>
FVH_HD void eig_edge_vjp_n(T g, const T* q, T eb, T* nb) {
    const T c = sqrt(pressure(g, q) * g / q[0]);
    nb[0] = nb[0] + eb * (c + sabs(q[1] / q[0]));
    nb[1] = nb[1] + eb * (c + sabs(q[2] / q[0]));
}

// d phi / d y of a limiter kind at y.
template <class T, int PHI>
FVH_HD T phi_deriv(T y, T yt) {
    return phi<Dual<T>, PHI>(Dual<T>(y, T(1)), Dual<T>(yt)).d;
}

}  // namespace fvh
