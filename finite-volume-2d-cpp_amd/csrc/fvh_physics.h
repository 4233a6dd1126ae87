// fvh_physics.h -- the per-edge, per-boundary and per-cell arithmetic that
// fvhyper leaves to user plugins, as host+device functions selected by kind.
//
// The same source serves the GPU kernels and the host (plugin probing in the
// drop-in layer).  With FP contraction off and IEEE division/sqrt on both
// sides, the two give identical bits, which is what lets the FP64 path match
// the reference CPU run exactly.  Expressions keep the association of the
// reference plugins on purpose: a reordering that is equal in exact
// arithmetic changes the last bits.
#pragma once

#include <math.h>

// Host and device by default; a host-only translation unit (a plain C++
// compiler, no HIP headers) defines FVH_HD as `inline` before including.
#ifndef FVH_HD
#include <hip/hip_runtime.h>
#define FVH_HD __host__ __device__ inline
#endif

namespace fvh {

// std::min / std::max as libstdc++ defines them (NaN and signed-zero
// behaviour included): min(a,b) = (b<a)?b:a, max(a,b) = (a<b)?b:a.
template <class T> FVH_HD T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> FVH_HD T smax(T a, T b) { return (a < b) ? b : a; }
template <class T> FVH_HD T sabs(T a) { return fabs(a); }

template <class T> struct Vec4 { T v[4]; };

// Smoothed kinks (FVH_FLAG_SMOOTH_KINKS; an extension with no reference
// counterpart, for derivative checks): with a width w > 0,
//   |x|_w = x^2 / sqrt(x^2 + w^2),  min_w(a, b) = (a + b - |a - b|_w) / 2,
//   max_w(a, b) = (a + b + |a - b|_w) / 2.
// Infinitely differentiable; |0|_w = 0 and min_w(a, a) = max_w(a, a) = a
// exactly, and |x|_w -> |x|, min_w -> min, max_w -> max for |x|, |a - b| >> w.
// min_w and max_w are symmetric bit for bit ((b-a)^2 rounds as (a-b)^2).
template <class T> FVH_HD T kabs(T x, T w) { return x * x / sqrt(x * x + w * w); }
template <class T> FVH_HD T kmin(T a, T b, T w) { return T(0.5) * (a + b - kabs(a - b, w)); }
template <class T> FVH_HD T kmax(T a, T b, T w) { return T(0.5) * (a + b + kabs(a - b, w)); }
// d|x|_w/dx = x (x^2 + 2 w^2) / (x^2 + w^2)^(3/2), in (-1, 1); the partial
// derivatives of min_w(a, b) are (1 - d)/2 and (1 + d)/2 with d = kabs_d(a - b),
// those of max_w(a, b) (1 + d)/2 and (1 - d)/2.
template <class T> FVH_HD T kabs_d(T x, T w) {
    const T s2 = x * x + w * w;
    return x * (x * x + T(2) * w * w) / (s2 * sqrt(s2));
}
// The limiter's sigma blend (explicit.cpp:226-238) over y = spread / (K sqrt A)^3 - 1
// in (0, 1]: the reference's cubic 2y^3 - 3y^2 + 1, once differentiable at
// both ends; with smoothed kinks 1 - S(y), S the quartic-order smoothstep
// y^5 (126 - 420y + 540y^2 - 315y^3 + 70y^4), four times differentiable at
// both ends (fine meshes make the blend narrow in state space: a jump in the
// second derivative there spoils difference quotients); and d sigma / dy.
template <class T> FVH_HD T sigma_cubic(T y) { return T(2) * y * y * y - T(3) * y * y + T(1); }
template <class T> FVH_HD T sigma_smooth(T y) {
    const T y2 = y * y;
    return T(1) - y2 * y2 * y * (T(126) + y * (T(-420) + y * (T(540) + y * (T(-315) + T(70) * y))));
}
template <class T> FVH_HD T sigma_smooth_d(T y) {
    const T y2 = y * y, z = T(1) - y, z2 = z * z;
    return T(-630) * (y2 * y2) * (z2 * z2);
}
// Venkatakrishnan's phi(d / q) = (r^2 + 2r) / (r^2 + r + 2) with r = d / q,
// written without the quotient and regularised by e2 > 0 as in his original
// form: (d^2 + 2 d q + e2) / (d^2 + d q + 2 q^2 + e2), smooth everywhere
// (value 1 at d = q = 0, where the quotient form is singular).
template <class T> FVH_HD T venkat_dq(T d, T q, T e2) {
    return (d * d + T(2) * d * q + e2) / (d * d + d * q + T(2) * q * q + e2);
}

// Pressure of an ideal gas, calc_p (examples/forwardStep/main.cpp:37-39).
template <class T>
FVH_HD T pressure(T g, const T* q) {
    return (g - T(1)) * (q[3] - T(0.5) / q[0] * (q[1] * q[1] + q[2] * q[2]));
}

// Roe flux-difference splitting with the "cures" entropy fix.
//  joined=false: examples/shockTube/main.cpp:68-135 (also naca/main.cpp:71-137)
//  joined=true : examples/forwardStep/main.cpp:72-135
// The two differ only in how the central flux is associated.
// SM: the smoothed entropy cure, |.| -> |.|_w, min/max -> min_w/max_w with
// the velocity width w (FVH_FLAG_SMOOTH_KINKS).  PRE: the caller passes the
// two face states' pressures (pressure(g, qi), pressure(g, qj), computed
// already) as pi_in, pj_in.
template <class T, bool JOINED, bool SM = false, bool PRE = false>
FVH_HD void roe_flux(T g, T* f, const T* qi, const T* qj, T n0, T n1, T w = T(0), T pi_in = T(0),
                     T pj_in = T(0)) {
    const T pi = PRE ? pi_in : pressure(g, qi);
    const T pj = PRE ? pj_in : pressure(g, qj);
    const T Vi = (qi[1] * n0 + qi[2] * n1) / qi[0];
    const T Vj = (qj[1] * n0 + qj[2] * n1) / qj[0];
    if (JOINED) {
        f[0] = (qi[0] * Vi + qj[0] * Vj) * T(0.5);
        f[1] = (qi[1] * Vi + pi * n0 + qj[1] * Vj + pj * n0) * T(0.5);
        f[2] = (qi[2] * Vi + pi * n1 + qj[2] * Vj + pj * n1) * T(0.5);
        f[3] = ((qi[3] + pi) * Vi + (qj[3] + pj) * Vj) * T(0.5);
    } else {
        f[0] = qi[0] * Vi;
        f[1] = qi[1] * Vi + pi * n0;
        f[2] = qi[2] * Vi + pi * n1;
        f[3] = (qi[3] + pi) * Vi;
        f[0] = f[0] + qj[0] * Vj;
        f[1] = f[1] + (qj[1] * Vj + pj * n0);
        f[2] = f[2] + (qj[2] * Vj + pj * n1);
        f[3] = f[3] + (qj[3] + pj) * Vj;
        f[0] = f[0] * T(0.5);
        f[1] = f[1] * T(0.5);
        f[2] = f[2] * T(0.5);
        f[3] = f[3] * T(0.5);
    }
    const T pL = pi, pR = pj;
    const T uL = qi[1] / qi[0];
    const T uR = qj[1] / qj[0];
    const T vL = qi[2] / qi[0];
    const T vR = qj[2] / qj[0];
    const T srhoL = sqrt(qi[0]);
    const T srhoR = sqrt(qj[0]);
    const T rho = srhoR * srhoL;
    const T den = srhoL + srhoR;
    const T u = (uL * srhoL + uR * srhoR) / den;
    const T v = (vL * srhoL + vR * srhoR) / den;
    const T h = ((qi[3] + pL) / qi[0] * srhoL + (qj[3] + pR) / qj[0] * srhoR) / den;
    const T q2 = u * u + v * v;
    const T c = sqrt((g - T(1)) * (h - T(0.5) * q2));
    const T V = u * n0 + v * n1;
    const T VR = uR * n0 + vR * n1;
    const T VL = uL * n0 + vL * n1;
    const T lcm = SM ? kabs(kmin(V - c, VL - c, w), w) : sabs(smin(V - c, VL - c));
    const T lc = SM ? kabs(V, w) : sabs(V);
    const T lcp = SM ? kabs(kmax(V + c, VR + c, w), w) : sabs(smax(V + c, VR + c));
    const T dp = pR - pL;
    const T dV = VR - VL;
    const T rc = rho * c;
    const T c2x2 = T(2) * c * c;
    const T kF1 = lcm * (dp - rc * dV) / c2x2;
    const T kF0 = lc * ((qj[0] - qi[0]) - dp / (c * c));
    const T kFr = lc * rho;
    const T kF5 = lcp * (dp + rc * dV) / c2x2;
    f[0] = f[0] - T(0.5) * (kF1 + kF0 + kF5);
    f[1] = f[1] - T(0.5) * (kF1 * (u - c * n0) + kF0 * u + kFr * (uR - uL - dV * n0) + kF5 * (u + c * n0));
    f[2] = f[2] - T(0.5) * (kF1 * (v - c * n1) + kF0 * v + kFr * (vR - vL - dV * n1) + kF5 * (v + c * n1));
    f[3] = f[3] - T(0.5) * (kF1 * (h - c * V) + kF0 * q2 * T(0.5) + kFr * (u * (uR - uL) + v * (vR - vL) - V * dV) + kF5 * (h + c * V));
}

// limiter_func kinds.
//  zero:      examples/shockTube/main.cpp:60-62
//  Michalak:  examples/forwardStep/main.cpp:57-66 (y_t a parameter, 2 there)
//  minmod:    examples/burgers/main.cpp:51-53
//  Venkatakrishnan: (r^2 + 2r)/(r^2 + r + 2); no reference example.
//  PHI_VENKAT_S: the Venkatakrishnan limiter with smoothed kinks (library
//  internal kind of FVH_FLAG_SMOOTH_KINKS; the limiter function itself is
//  venkat_dq, see limit_cell).
enum PhiKind { PHI_ZERO = 0, PHI_MICHALAK = 1, PHI_VENKAT = 2, PHI_MINMOD = 3, PHI_VENKAT_S = 4 };

template <class T, int PHI>
FVH_HD T phi(T y, T yt) {
    if (PHI == PHI_ZERO) return T(0);
    if (PHI == PHI_MICHALAK) {
        if (y >= yt) return T(1);
        const T a = T(1) / (yt * yt) - T(2) / (yt * yt * yt);
        const T b = T(-1.5) * a * yt - T(0.5) / yt;
        return a * y * y * y + b * y * y + y;
    }
    if (PHI == PHI_VENKAT || PHI == PHI_VENKAT_S) return (y * y + T(2) * y) / (y * y + y + T(2));
    return smax(T(0), smin(T(1), y));
}

// Boundary-condition kinds.  b receives the ghost state, q is the interior
// face state, (n0, n1) the outward unit normal.
//  copy:    examples/shockTube/main.cpp:158-163
//  mirror:  examples/forwardStep/main.cpp:262-269 (= naca/main.cpp:260-267)
//  inlet/outlet characteristic: examples/forwardStep/main.cpp:194-261
//  farfield characteristic:     examples/naca/main.cpp:192-259
// Both characteristic kinds keep the reference's b[2] = b[1]*(...) spelling
// in their subsonic branches; the freestream (rho, u, v, p) is a parameter.
//  no-slip wall: examples/flatPlate/main.cpp:361-366
//  farfield, b[0] spelling: examples/flatPlate/main.cpp:286-353 (freestream a
//  parameter; its subsonic outlet keeps the reference's v_a)
enum BcKind { BC_COPY = 0, BC_MIRROR = 1, BC_INLET_OUTLET = 2, BC_FARFIELD = 3, BC_NOSLIP = 4, BC_FARFIELD_V = 5 };

template <class T>
struct BcParam {
    int kind;
    T rho, u, v, p;
};

template <class T>
FVH_HD void boundary(const BcParam<T>& P, T g, T* b, const T* q, T n0, T n1) {
    if (P.kind == BC_COPY) {
        b[0] = q[0]; b[1] = q[1]; b[2] = q[2]; b[3] = q[3];
        return;
    }
    if (P.kind == BC_MIRROR) {
        const T un = n0 * q[1] + n1 * q[2];
        b[0] = q[0];
        b[1] = q[1] - T(2) * n0 * un;
        b[2] = q[2] - T(2) * n1 * un;
        b[3] = q[3];
        return;
    }
    if (P.kind == BC_NOSLIP) {
        b[0] = q[0];
        b[1] = -q[1];
        b[2] = -q[2];
        b[3] = q[3];
        return;
    }
    T bv[4];
    bv[0] = P.rho;
    bv[1] = bv[0] * P.u;
    bv[2] = bv[0] * P.v;
    bv[3] = P.p / (g - T(1)) + T(0.5) / bv[0] * (bv[1] * bv[1] + bv[2] * bv[2]);
    const T U0 = q[1] / q[0];
    const T U1 = q[2] / q[0];
    const T u_norm = sqrt(U0 * U0 + U1 * U1);
    const T u_dot_n = U0 * n0 + U1 * n1;
    const T p = pressure(g, q);
    const T c = sqrt(g * p / q[0]);
    const T mach = u_norm / c;
    if (mach > T(1)) {
        if (u_dot_n < T(0)) {
            b[0] = bv[0]; b[1] = bv[1]; b[2] = bv[2]; b[3] = bv[3];
        } else {
            b[0] = q[0]; b[1] = q[1]; b[2] = q[2]; b[3] = q[3];
        }
        return;
    }
    const T pa = pressure(g, bv);
    const T rhoa = bv[0];
    const T ua = bv[1] / bv[0];
    const T va = bv[2] / bv[0];
    const T pd = p;
    const T rhod = q[0];
    const T ud = q[1] / q[0];
    const T vd = q[2] / q[0];
    const T rho0 = q[0];
    const T c0 = c;
    if (u_dot_n < T(0)) {
        if (P.kind == BC_INLET_OUTLET) {
            const T pb = pd;
            b[0] = bv[0]; b[1] = bv[1]; b[2] = bv[2];
            b[3] = pb / (g - T(1)) + T(0.5) / b[0] * (b[1] * b[1] + b[2] * b[2]);
        } else {
            const T pb = T(0.5) * (pa + pd - rho0 * c0 * (n0 * (ua - ud) + n1 * (va - vd)));
            b[0] = rhoa + (pb - pa) / (c0 * c0);
            b[1] = b[0] * (ua - n0 * (pa - pb) / (rho0 * c0));
            b[2] = (P.kind == BC_FARFIELD_V ? b[0] : b[1]) * (va - n1 * (pa - pb) / (rho0 * c0));
            b[3] = pb / (g - T(1)) + T(0.5) / b[0] * (b[1] * b[1] + b[2] * b[2]);
        }
    } else {
        const T pb = pa;
        b[0] = rhod + (pb - pd) / (c0 * c0);
        b[1] = b[0] * (ud + n0 * (pd - pb) / (rho0 * c0));
        b[2] = (P.kind == BC_FARFIELD_V ? b[0] : b[1]) * (va + n1 * (pd - pb) / (rho0 * c0));
        b[3] = pb / (g - T(1)) + T(0.5) / b[0] * (b[1] * b[1] + b[2] * b[2]);
    }
}

// Per-edge wave-speed terms of the calc_dt kinds.
//  summed CFL (forwardStep/main.cpp:167-171): c + |u.n|
// eig_cell is the part that depends on the state only (sound speed and
// velocity), so a cell's is computed once for all its edges:
// eig_sum(eig_cell(g, q), n0, n1) is eig_sum(g, q, n0, n1) bit for bit.
template <class T>
struct EigCell {
    T c, u, v;
};
template <class T>
FVH_HD EigCell<T> eig_cell(T g, const T* q) {
    return EigCell<T>{sqrt(pressure(g, q) * g / q[0]), q[1] / q[0], q[2] / q[0]};
}
template <class T>
FVH_HD T eig_sum(const EigCell<T>& e, T n0, T n1) {
    return e.c + sabs(e.u * n0 + e.v * n1);
}
template <class T>
FVH_HD T eig_sum(T g, const T* q, T n0, T n1) {
    return eig_sum(eig_cell(g, q), n0, n1);
}
// the smoothed form (FVH_FLAG_SMOOTH_KINKS): c + |u.n|_w
template <class T>
FVH_HD T eig_sum_k(const EigCell<T>& e, T n0, T n1, T w) {
    return e.c + kabs(e.u * n0 + e.v * n1, w);
}
template <class T>
FVH_HD T eig_sum_k(T g, const T* q, T n0, T n1, T w) {
    return eig_sum_k(eig_cell(g, q), n0, n1, w);
}
//  edge-minimum local (naca/main.cpp:169-173): (c+|u|) n0 + (c+|v|) n1
template <class T>
FVH_HD T eig_edge(T g, const T* q, T n0, T n1) {
    const T c = sqrt(pressure(g, q) * g / q[0]);
    return (c + sabs(q[1] / q[0])) * n0 + (c + sabs(q[2] / q[0])) * n1;
}

//  flatPlate local form (flatPlate/main.cpp:261-262): (c + |n0 u| + |n1 v|)
template <class T>
FVH_HD T eig_abs(T g, const T* q, T n0, T n1) {
    const T c = sqrt(pressure(g, q) * g / q[0]);
    return c + sabs(n0 * q[1] / q[0]) + sabs(n1 * q[2] / q[0]);
}
//  viscous spectral term of both viscous forms (flatPlate/main.cpp:264-267):
//  max(4/(3 rho), gamma/rho) (mu/pr) len^2 / A
template <class T>
FVH_HD T eig_visc(T g, T mu, T pr, const T* q, T len, T A) {
    return smax(T(4) / (T(3) * q[0]), g / q[0]) * (mu / pr) * len * len / A;
}

// gradient_for_diffusion (explicit.cpp:70-109): the mean of the two cell
// gradients with its component along the centre line replaced by the
// difference quotient of the cell states.
template <class T>
FVH_HD void face_gradient(T* gxf, T* gyf, const T* gxi, const T* gyi, const T* gxj, const T* gyj, const T* qi,
                          const T* qj, T cix, T ciy, T cjx, T cjy) {
    T t0 = cix - cjx;
    T t1 = ciy - cjy;
    const T l = sqrt(t0 * t0 + t1 * t1);
    t0 = t0 / l;
    t1 = t1 / l;
    for (int k = 0; k < 4; ++k) {
        const T gd = (qi[k] - qj[k]) / l;
        const T gxb = (gxi[k] + gxj[k]) * T(0.5);
        const T gyb = (gyi[k] + gyj[k]) * T(0.5);
        const T dot = gxb * t0 + gyb * t1;
        gxf[k] = gxb - (dot - gd) * t0;
        gyf[k] = gyb - (dot - gd) * t1;
    }
}

// Laminar viscous terms subtracted from the Roe flux (flatPlate/main.cpp:
// 193-224 with calc_grad_p/t/u/v of :43-92): the stresses and heat flux of
// the mean face state qc = (qi + qj)/2 and the conservative face gradient
// (gx, gy).
template <class T>
FVH_HD void viscous_terms(T g, T mu, T pr, T cp, T r, T* f, const T* qi, const T* qj, const T* gx, const T* gy,
                          T n0, T n1) {
    T qc[4];
    for (int k = 0; k < 4; ++k) qc[k] = T(0.5) * (qi[k] + qj[k]);
    const T q00 = qc[0] * qc[0];
    // pressure gradient (calc_grad_p)
    T gp0 = T(2) * gx[3];
    gp0 = gp0 - (gx[1] * qc[1] / qc[0] + qc[1] * (gx[1] * qc[0] - gx[0] * qc[1]) / q00);
    gp0 = gp0 - (gx[2] * qc[2] / qc[0] + qc[2] * (gx[2] * qc[0] - gx[0] * qc[2]) / q00);
    gp0 = gp0 * (T(0.5) * (g - T(1)));
    T gp1 = T(2) * gy[3];
    gp1 = gp1 - (gy[1] * qc[1] / qc[0] + qc[1] * (gy[1] * qc[0] - gy[0] * qc[1]) / q00);
    gp1 = gp1 - (gy[2] * qc[2] / qc[0] + qc[2] * (gy[2] * qc[0] - gy[0] * qc[2]) / q00);
    gp1 = gp1 * (T(0.5) * (g - T(1)));
    // temperature gradient (calc_grad_t)
    const T p = pressure(g, qc);
    const T gt0 = (T(1) / r) * ((gp0 * qc[0] - gx[0] * p) / q00);
    const T gt1 = (T(1) / r) * ((gp1 * qc[0] - gy[0] * p) / q00);
    // velocity gradients (calc_grad_u, calc_grad_v)
    const T gu0 = (qc[0] * gx[1] - qc[1] * gx[0]) / q00;
    const T gu1 = (qc[0] * gy[1] - qc[1] * gy[0]) / q00;
    const T gv0 = (qc[0] * gx[2] - qc[2] * gx[0]) / q00;
    const T gv1 = (qc[0] * gy[2] - qc[2] * gy[0]) / q00;
    const T div_v = gu0 + gv1;
    const T txx = T(2) * mu * (gu0 - div_v / T(3));
    const T tyy = T(2) * mu * (gv1 - div_v / T(3));
    const T txy = mu * (gu1 + gv0);
    const T k = cp * mu / pr;
    const T ph0 = qc[1] / qc[0] * txx + qc[2] / qc[0] * txy + k * gt0;
    const T ph1 = qc[1] / qc[0] * txy + qc[2] / qc[0] * tyy + k * gt1;
    f[1] = f[1] - (n0 * txx + n1 * txy);
    f[2] = f[2] - (n0 * txy + n1 * tyy);
    f[3] = f[3] - (n0 * ph0 + n1 * ph1);
}

}  // namespace fvh
