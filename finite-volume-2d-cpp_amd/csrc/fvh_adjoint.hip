// fvh_adjoint.hip -- reverse-mode (discrete adjoint) kernels of one RK stage
// of the explicit path, and of its time step, initial boundary update and
// boundary-pressure objective.
//
// A stage maps its input state qin (all cells) to its output (real cells:
// q0 + qt dt alpha; ghosts: boundary functions of the updated interior face
// states).  Its reverse, given the adjoint of the output, runs four passes:
//   update  (per real cell)  boundary functions reversed into the cell and
//                            its boundary increments; adjoint of q0, qt, dt
//   flux    (per edge)       adjoint of qt to the edge flux, reversed
//                            through the Roe flux into both face states
//   gradlim (per real cell)  face states reversed into the cell state, its
//                            gradient and its limiter; limiter and gradient
//                            reversed; contributions to neighbours written
//                            per half-edge
//   gather  (per cell)       own terms plus the neighbours' half-edge
//                            contributions, in CSR order
// No atomics: every sum is a gather in a fixed order, so the adjoint is
// deterministic.  Forward quantities (stage states, gradients, limiters, qt)
// come from a replay of the step; comparisons and branches (min/max, |.|,
// limiter ranges, positivity fallback, boundary regimes) are re-evaluated on
// them and differentiated piecewise.
#include "fvh_adjoint.h"
#include "fvh_device.h"

#include <algorithm>

namespace fvh {
namespace {

template <class T>
__device__ inline void q4_add(T* a, const Q4<T>& b) {
#pragma unroll
    for (int k = 0; k < 4; ++k) a[k] = a[k] + b.v[k];
}

// Adjoint of a stage's output state: an explicit array, or gathered on the
// fly from the next stage's reverse (its per-cell direct terms, per-half-edge
// face adjoints and per-half-edge neighbour contributions), in the order the
// gather pass uses.
template <class T>
struct LamSrc {
    const Q4<T>* lam;      // explicit, or nullptr: gathered from the three below
    const Q4<T>* direct;
    const Q4<T>* face;
    const Q4<T>* nb;
};

template <class T>
__device__ inline Q4<T> lam_of(const MeshView<T>& m, const LamSrc<T>& L, int c) {
    if (L.lam) return L.lam[c];
    T v[4] = {T(0), T(0), T(0), T(0)};
    const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
    if (c < m.n_real) {
        q4_add(v, L.direct[c]);
    } else {
        for (int h = h0; h < h1; ++h) q4_add(v, L.face[h]);   // a ghost's face state is the ghost state
    }
    for (int h = h0; h < h1; ++h)
        if (m.he_nbr[h] < m.n_real) q4_add(v, L.nb[m.he_twin[h]]);
    return Q4<T>{{v[0], v[1], v[2], v[3]}};
}

// ---- boundary pass (per ghost cell) -----------------------------------------
// ghost = boundary(interior face state of the owner, n): the adjoint of the
// ghost's output reversed into the owner's updated state (lam_bq) and, with
// reconstruction, into the owner's face increment (lam_bincr).  kBcLanes
// lanes per ghost, one Jacobian column each (lanes 0-3 the state, 4-5 the
// normal when the geometry adjoint is formed), combined by shuffles.
constexpr int kBcLanes = 8;

template <class T, bool LIN, bool GEO>
__global__ __launch_bounds__(kBlock) void k_adj_bc(MeshView<T> m, PhysParams pp, const BcParam<T>* __restrict__ bcs,
                                                   const Q4<T>* __restrict__ qout, const Q4<T>* __restrict__ gx,
                                                   const Q4<T>* __restrict__ gy, const Q4<T>* __restrict__ lm,
                                                   LamSrc<T> L, Q4<T>* __restrict__ lam_bq,
                                                   Q4<T>* __restrict__ lam_bincr, GeoAdj<T> ga) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int lane = t % kBcLanes;
    const int gc = m.n_real + t / kBcLanes;
    // every lane of a group takes the same branches; a group never straddles a wavefront
    const bool live = gc < m.n_total;
    const int hg = live ? m.he_off[gc] : 0;
    const bool has_edge = live && m.he_off[gc + 1] > hg;   // (a repeated boundary element leaves a ghost with no edge)
    T col = T(0);
    bool fallback = !LIN;
    int h = 0;
    if (has_edge) {
        h = m.he_twin[hg];   // the owner's half-edge
        const int c = m.he_nbr[hg];
        const uint32_t ue = m.he_edge[h];
        const Q4<T> geo = m.e_geo[ue & kEdgeMask];
        const T g = T(pp.gamma);
        const Q4<T> out = qout[c];
        T qi[4];
        if (LIN) {
            T ax[4], ay[4], al[4];
            const Q4<T> a = gx[c], b = gy[c], d = lm[c];
#pragma unroll
            for (int k = 0; k < 4; ++k) { ax[k] = a.v[k]; ay[k] = b.v[k]; al[k] = d.v[k]; }
            const Q4<T> inc = face_inc(ax, ay, al, m.e_c[ue & kEdgeMask], m.c_c[c]);
#pragma unroll
            for (int k = 0; k < 4; ++k) qi[k] = out.v[k] + inc.v[k];
            if (pp.positivity && !physical(g, qi)) {
                fallback = true;
#pragma unroll
                for (int k = 0; k < 4; ++k) qi[k] = out.v[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) qi[k] = out.v[k];
        }
        const Q4<T> lg = lam_of(m, L, gc);
        if (lane < 4 || (GEO && lane < 6)) col = boundary_dir(bcs[m.he_bc[h]], g, qi, geo.v[0], geo.v[1], lane, lg.v);
    }
    const int base = (threadIdx.x / kBcLanes) * kBcLanes;
    T v[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) v[j] = __shfl(col, base + j);
    if (!live || lane != 0) return;
    // (0 + x: the sum boundary_vjp forms for one edge)
    const Q4<T> qb4{{T(0) + v[0], T(0) + v[1], T(0) + v[2], T(0) + v[3]}};
    lam_bq[gc - m.n_real] = qb4;
    lam_bincr[gc - m.n_real] = fallback ? Q4<T>{{T(0), T(0), T(0), T(0)}} : qb4;
    if (GEO && has_edge) {
        Q4<T> a = ga.he[h];
        a.v[0] = a.v[0] + v[4];
        a.v[1] = a.v[1] + v[5];
        ga.he[h] = a;
    }
}

// ---- update pass --------------------------------------------------------
template <class T, bool GEO>
__global__ __launch_bounds__(kBlock) void k_adj_update(MeshView<T> m, TileMap tm, PhysParams pp,
                                                       const Q4<T>* __restrict__ qt, const T* __restrict__ dt,
                                                       const T* __restrict__ dt_g, T alpha, LamSrc<T> L,
                                                       const Q4<T>* __restrict__ lam_bq, Q4<T>* __restrict__ lam_q0,
                                                       Q4<T>* __restrict__ lam_qt, T* __restrict__ lam_dt,
                                                       double* __restrict__ ld_blk, GeoAdj<T> ga) {
    const int lb = xcd_block(tm.n_rt);
    if (lb < 0) return;
    const int c = tm.c0[lb] + threadIdx.x;
    double ld = 0.0;
    if (c < tm.c0[lb + 1]) {
        T lam[4];
        const Q4<T> lo = lam_of(m, L, c);
#pragma unroll
        for (int k = 0; k < 4; ++k) lam[k] = lo.v[k];
        // ghosts updated from this cell's face states (boundary pass)
        for (int h = m.he_off[c]; h < m.he_off[c + 1]; ++h) {
            if (!(m.he_edge[h] & kBcBit)) continue;
            q4_add(lam, lam_bq[m.he_nbr[h] - m.n_real]);
        }
        Q4<T> a0 = lam_q0[c];
#pragma unroll
        for (int k = 0; k < 4; ++k) a0.v[k] = a0.v[k] + lam[k];
        lam_q0[c] = a0;
        const T d = dt_g ? dt_g[0] : dt[c];
        const Q4<T> t = qt[c];
        Q4<T> lq;
        T s = T(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lq.v[k] = lam[k] * d * alpha;
            s = s + lam[k] * t.v[k] * alpha;
        }
        lam_qt[c] = lq;
        if (GEO) {
            // qt = (sum of signed f len) / A, so d qt / d A = -qt / A
            const T A = m.c_area[c];
            T sA = T(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) sA = sA - lq.v[k] * t.v[k] / A;
            Q4<T> a = ga.c[c];
            a.v[0] = a.v[0] + sA;
            ga.c[c] = a;
        }
        if (dt_g) ld = (double)s;
        else if (lam_dt) lam_dt[c] = lam_dt[c] + s;
    }
    if (ld_blk) {
        double v[4] = {ld, 0.0, 0.0, 0.0};
        block_sum4(v);
        if (threadIdx.x == 0) ld_blk[lb] = v[0];
    }
}

// ---- flux pass (one thread per edge) -------------------------------------
// SM: the smoothed entropy cure (FVH_FLAG_SMOOTH_KINKS)
template <class T, bool LIN, bool GEO, bool SM>
__global__ __launch_bounds__(kBlock) void k_adj_flux(MeshView<T> m, PhysParams pp, int n_edges,
                                                     const int4* __restrict__ epair, const Q4<T>* __restrict__ qin,
                                                     const Q4<T>* __restrict__ gx, const Q4<T>* __restrict__ gy,
                                                     const Q4<T>* __restrict__ lm, const Q4<T>* __restrict__ lam_qt,
                                                     Q4<T>* __restrict__ lam_face, GeoAdj<T> ga) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= n_edges) return;
    const int4 ep = epair[e];   // (half-edge side 0, half-edge side 1, cell side 0, cell side 1)
    const Q4<T> geo = m.e_geo[e];
    const G2<T> ec = m.e_c[e];
    const T g = T(pp.gamma);
    T fb[4] = {T(0), T(0), T(0), T(0)};
    if (ep.z < m.n_real) {
        const Q4<T> l = lam_qt[ep.z];
        const T w = geo.v[2] * m.c_inv_area[ep.z];
#pragma unroll
        for (int k = 0; k < 4; ++k) fb[k] = fb[k] - l.v[k] * w;
    }
    if (ep.w < m.n_real) {
        const Q4<T> l = lam_qt[ep.w];
        const T w = geo.v[2] * m.c_inv_area[ep.w];
#pragma unroll
        for (int k = 0; k < 4; ++k) fb[k] = fb[k] + l.v[k] * w;
    }
    T fs[2][4];
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const int c = sd ? ep.w : ep.z;
        const Q4<T> q = qin[c];
        if (LIN) {
            T ax[4], ay[4], al[4];
            const Q4<T> a = gx[c], b = gy[c], d = lm[c];
#pragma unroll
            for (int k = 0; k < 4; ++k) { ax[k] = a.v[k]; ay[k] = b.v[k]; al[k] = d.v[k]; }
            const Q4<T> inc = face_inc(ax, ay, al, ec, m.c_c[c]);
#pragma unroll
            for (int k = 0; k < 4; ++k) fs[sd][k] = q.v[k] + inc.v[k];
            if (pp.positivity && !physical(g, fs[sd]))
#pragma unroll
                for (int k = 0; k < 4; ++k) fs[sd][k] = q.v[k];
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) fs[sd][k] = q.v[k];
        }
    }
    T qib[4] = {T(0), T(0), T(0), T(0)}, qjb[4] = {T(0), T(0), T(0), T(0)}, nb[2] = {T(0), T(0)};
    roe_flux_vjp<T, SM>(g, fs[0], fs[1], geo.v[0], geo.v[1], fb, qib, qjb, nb, T(pp.kwv));
    Q4<T> a, b;
#pragma unroll
    for (int k = 0; k < 4; ++k) { a.v[k] = qib[k]; b.v[k] = qjb[k]; }
    lam_face[ep.x] = a;
    lam_face[ep.y] = b;
    if (GEO) {
        // qt_i -= f len / A_i, qt_j += f len / A_j: the length's adjoint is
        // sum_k f_k (lam_qt_j / A_j - lam_qt_i / A_i)_k
        T f[4];
        roe_flux<T, false, SM>(g, f, fs[0], fs[1], geo.v[0], geo.v[1], T(pp.kwv));
        T ll = T(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) ll = ll + fb[k] / geo.v[2] * f[k];
        Q4<T> ge = ga.e[e];
        ge.v[0] = ge.v[0] + nb[0];
        ge.v[1] = ge.v[1] + nb[1];
        ge.v[2] = ge.v[2] + ll;
        ga.e[e] = ge;
    }
}

// ---- gradient/limiter pass -------------------------------------------------
template <class T, int PHI, bool LIN, bool GEO>
__global__ __launch_bounds__(kBlock) void k_adj_gradlim(MeshView<T> m, PhysParams pp, int grad,
                                                        int lim_on, const Q4<T>* __restrict__ qin,
                                                        const Q4<T>* __restrict__ gxs, const Q4<T>* __restrict__ gys,
                                                        const Q4<T>* __restrict__ lms,
                                                        const Q4<T>* __restrict__ lam_face,
                                                        const Q4<T>* __restrict__ lam_bincr,
                                                        Q4<T>* __restrict__ lam_direct,
                                                        Q4<T>* __restrict__ nb_contrib, GeoAdj<T> ga) {
    // one lane per real cell over the whole range (nothing here is per
    // tile, and tiles of fewer than kBlock cells would leave lanes idle)
    const int lb = xcd_block((m.n_real + kBlock - 1) / kBlock);
    if (lb < 0) return;
    const int c = lb * kBlock + threadIdx.x;
    if (c >= m.n_real) return;
    const T g = T(pp.gamma);
    const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
    const Q4<T> qc = qin[c];
    const G2<T> cc = m.c_c[c];
    T gx[4], gy[4], lim[4];
    {
        const Q4<T> a = gxs[c], b = gys[c], d = lms[c];
#pragma unroll
        for (int k = 0; k < 4; ++k) { gx[k] = a.v[k]; gy[k] = b.v[k]; lim[k] = d.v[k]; }
    }
    T lqc[4] = {T(0), T(0), T(0), T(0)};
    T lgx[4] = {T(0), T(0), T(0), T(0)}, lgy[4] = {T(0), T(0), T(0), T(0)}, llim[4] = {T(0), T(0), T(0), T(0)};
    T lA = T(0), lcx = T(0), lcy = T(0);   // geometry adjoint of this cell (GEO)
    // contributions to the neighbours, per half-edge (a real cell has at most
    // kMaxHe; the host checks), stored once at the end
    T nbc[kMaxHe][4];
#pragma unroll
    for (int j = 0; j < kMaxHe; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) nbc[j][k] = T(0);
    // face states q + inc (and boundary increments) reversed
    for (int h = h0; h < h1; ++h) {
        const uint32_t ue = m.he_edge[h];
        const G2<T> ec = m.e_c[ue & kEdgeMask];
        const Q4<T> lf = lam_face[h];
        q4_add(lqc, lf);
        if (!LIN) continue;
        const Q4<T> inc = face_inc(gx, gy, lim, ec, cc);
        T fs[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) fs[k] = qc.v[k] + inc.v[k];
        T li[4];
        const bool fallback = pp.positivity && !physical(g, fs);
#pragma unroll
        for (int k = 0; k < 4; ++k) li[k] = fallback ? T(0) : lf.v[k];
        if (ue & kBcBit) {
            const Q4<T> lb2 = lam_bincr[m.he_nbr[h] - m.n_real];
#pragma unroll
            for (int k = 0; k < 4; ++k) li[k] = li[k] + lb2.v[k];
        }
        const T dx = ec.v[0] - cc.v[0];
        const T dy = ec.v[1] - cc.v[1];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lgx[k] = lgx[k] + li[k] * lim[k] * dx;
            lgy[k] = lgy[k] + li[k] * lim[k] * dy;
            llim[k] = llim[k] + li[k] * (gx[k] * dx + gy[k] * dy);
        }
        if (GEO) {
            // inc = (gx dx + gy dy) lim with (dx, dy) = edge centre - cell centre
            T ldx = T(0), ldy = T(0);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ldx = ldx + li[k] * gx[k] * lim[k];
                ldy = ldy + li[k] * gy[k] * lim[k];
            }
            G2<T> a = ga.hc[h];
            a.v[0] = a.v[0] + ldx;
            a.v[1] = a.v[1] + ldy;
            ga.hc[h] = a;
            lcx = lcx - ldx;
            lcy = lcy - ldy;
        }
    }
    // limiter reversed (explicit.cpp:174-258)
    if (lim_on && PHI == PHI_VENKAT_S) {
        // smoothed kinks: every step of the min_w / max_w chains over the
        // neighbours, of the blend on the sign of dqg and of the running
        // min_w over the half-edges has a derivative; each is reversed
        const int nh = h1 - h0;
        const T K3a = [&] { const T Ka = T(pp.K) * m.c_sqrt_area[c]; return Ka * Ka * Ka; }();
        const T wl = T(pp.kwl);
        T qn[kMaxHe][4];
        T dxs[kMaxHe], dys[kMaxHe];
#pragma unroll
        for (int j = 0; j < kMaxHe; ++j) {
            const int h = min(h0 + j, h1 - 1);
            const Q4<T> a = qin[m.he_nbr[h]];
#pragma unroll
            for (int k = 0; k < 4; ++k) qn[j][k] = a.v[k];
            const G2<T> ec = m.e_c[m.he_edge[h] & kEdgeMask];
            dxs[j] = ec.v[0] - cc.v[0];
            dys[j] = ec.v[1] - cc.v[1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const T w = T(pp.kw[k]);
            // forward: the extrema chains a (min_w) and b (max_w)
            T a[kMaxHe + 1], b[kMaxHe + 1];
            a[0] = qc.v[k];
            b[0] = qc.v[k];
#pragma unroll
            for (int j = 0; j < kMaxHe; ++j) {
                a[j + 1] = j < nh ? kmin(a[j], qn[j][k], w) : a[j];
                b[j + 1] = j < nh ? kmax(b[j], qn[j][k], w) : b[j];
            }
            const T dmax = b[nh] - qc.v[k];
            const T dmin = a[nh] - qc.v[k];
            const T spread = (dmax - dmin) * (dmax - dmin);
            T sig, dsig = T(0);
            if (spread <= K3a) sig = T(1);
            else if (spread <= T(2) * K3a) {
                const T y = (spread / K3a - T(1));
                sig = sigma_smooth(y);
                dsig = sigma_smooth_d(y) / K3a;   // d sig / d spread
            } else sig = T(0);
            // forward: per half-edge limiter values and the running minimum
            T lv[kMaxHe], Lv[kMaxHe], run[kMaxHe + 1];
            run[0] = T(1);
#pragma unroll
            for (int j = 0; j < kMaxHe; ++j) {
                const T dqg = gx[k] * dxs[j] + gy[k] * dys[j];
                T l = T(1);
                if (sig < T(1)) {
                    const T t = dqg / sqrt(dqg * dqg + w * w);
                    const T d = T(0.5) * (dmax + dmin) + T(0.5) * (dmax - dmin) * t;
                    l = venkat_dq(d, dqg, w * w);
                }
                lv[j] = l;
                Lv[j] = sig + (T(1) - sig) * l;
                run[j + 1] = j < nh ? kmin(run[j], Lv[j], wl) : run[j];
            }
            // reverse of the running minimum
            T rb = llim[k];
            T Lb[kMaxHe];
#pragma unroll
            for (int j = kMaxHe - 1; j >= 0; --j) {
                Lb[j] = T(0);
                if (j >= nh) continue;
                const T dk = kabs_d(run[j] - Lv[j], wl);
                Lb[j] = rb * T(0.5) * (T(1) + dk);
                rb = rb * T(0.5) * (T(1) - dk);
            }
            // reverse of L = sig + (1 - sig) l and of l(d, dqg)
            T sigb = T(0), dmaxb = T(0), dminb = T(0);
#pragma unroll
            for (int j = 0; j < kMaxHe; ++j) {
                if (j >= nh) continue;
                sigb = sigb + Lb[j] * (T(1) - lv[j]);
                if (!(sig < T(1))) continue;
                const T lb = Lb[j] * (T(1) - sig);
                const T dqg = gx[k] * dxs[j] + gy[k] * dys[j];
                const T s2 = dqg * dqg + w * w;
                const T t = dqg / sqrt(s2);
                const T d = T(0.5) * (dmax + dmin) + T(0.5) * (dmax - dmin) * t;
                const T den = d * d + d * dqg + T(2) * dqg * dqg + w * w;
                const T num = d * d + T(2) * d * dqg + w * w;
                const T db = lb * ((T(2) * d + T(2) * dqg) * den - num * (T(2) * d + dqg)) / (den * den);
                T qgb = lb * (T(2) * d * den - num * (d + T(4) * dqg)) / (den * den);
                dmaxb = dmaxb + db * T(0.5) * (T(1) + t);
                dminb = dminb + db * T(0.5) * (T(1) - t);
                const T tb = db * T(0.5) * (dmax - dmin);
                qgb = qgb + tb * w * w / (s2 * sqrt(s2));
                lgx[k] = lgx[k] + qgb * dxs[j];
                lgy[k] = lgy[k] + qgb * dys[j];
                if (GEO) {
                    // dqg = gx dx + gy dy, (dx, dy) = edge centre - cell centre
                    G2<T> e2 = ga.hc[h0 + j];
                    e2.v[0] = e2.v[0] + qgb * gx[k];
                    e2.v[1] = e2.v[1] + qgb * gy[k];
                    ga.hc[h0 + j] = e2;
                    lcx = lcx - qgb * gx[k];
                    lcy = lcy - qgb * gy[k];
                }
            }
            // sigma(spread / K3a)
            const T lspread = sigb * dsig;
            if (GEO) lA = lA - sigb * dsig * spread / K3a * (T(1.5) * K3a / m.c_area[c]);
            dmaxb = dmaxb + T(2) * (dmax - dmin) * lspread;
            dminb = dminb - T(2) * (dmax - dmin) * lspread;
            lqc[k] = lqc[k] - dmaxb - dminb;
            // reverse of the extrema chains
            T ab = dminb, bb = dmaxb;
#pragma unroll
            for (int j = kMaxHe - 1; j >= 0; --j) {
                if (j >= nh) continue;
                const T da = kabs_d(a[j] - qn[j][k], w);
                nbc[j][k] = nbc[j][k] + ab * T(0.5) * (T(1) + da);
                ab = ab * T(0.5) * (T(1) - da);
                const T dbx = kabs_d(b[j] - qn[j][k], w);
                nbc[j][k] = nbc[j][k] + bb * T(0.5) * (T(1) - dbx);
                bb = bb * T(0.5) * (T(1) + dbx);
            }
            lqc[k] = lqc[k] + ab + bb;
        }
    } else if (lim_on) {
        // neighbour extrema and the half-edge that sets each (-1: the cell itself)
        T qmn[4], qmx[4];
        int hmn[4], hmx[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) { qmn[k] = qc.v[k]; qmx[k] = qc.v[k]; hmn[k] = -1; hmx[k] = -1; }
        for (int h = h0; h < h1; ++h) {
            const Q4<T> qn = qin[m.he_nbr[h]];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (qn.v[k] < qmn[k]) { qmn[k] = qn.v[k]; hmn[k] = h; }
                if (qmx[k] < qn.v[k]) { qmx[k] = qn.v[k]; hmx[k] = h; }
            }
        }
        const T Ka = T(pp.K) * m.c_sqrt_area[c];
        const T K3a = Ka * Ka * Ka;
        const T yt = T(pp.phi_yt);
        const T tol = T(1e-15);
        // the half-edge each variable's running minimum ends on (one pass, as the forward)
        T run[4] = {T(1), T(1), T(1), T(1)};
        int hsel[4] = {-1, -1, -1, -1};
        const bool any = llim[0] != T(0) || llim[1] != T(0) || llim[2] != T(0) || llim[3] != T(0);
        for (int h = h0; any && h < h1; ++h) {
            const G2<T> ec = m.e_c[m.he_edge[h] & kEdgeMask];
            const T dx = ec.v[0] - cc.v[0];
            const T dy = ec.v[1] - cc.v[1];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const T dqg = gx[k] * dx + gy[k] * dy;
                const T dmax = qmx[k] - qc.v[k];
                const T dmin = qmn[k] - qc.v[k];
                const T spread = (dmax - dmin) * (dmax - dmin);
                T sig;
                if (spread <= K3a) sig = T(1);
                else if (spread <= T(2) * K3a) {
                    const T y = (spread / K3a - T(1));
                    sig = T(2) * y * y * y - T(3) * y * y + T(1);
                } else sig = T(0);
                T l = T(1);
                if (sig < T(1)) {
                    if (dqg > tol) l = phi<T, PHI>(dmax / dqg, yt);
                    else if (dqg < -tol) l = phi<T, PHI>(dmin / dqg, yt);
                }
                l = sig + (T(1) - sig) * l;
                if (l < run[k]) { run[k] = l; hsel[k] = h; }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int hs = hsel[k];
            if (llim[k] == T(0)) continue;
            if (hs < 0) continue;
            const G2<T> ec = m.e_c[m.he_edge[hs] & kEdgeMask];
            const T dx = ec.v[0] - cc.v[0];
            const T dy = ec.v[1] - cc.v[1];
            const T dqg = gx[k] * dx + gy[k] * dy;
            const T dmax = qmx[k] - qc.v[k];
            const T dmin = qmn[k] - qc.v[k];
            const T spread = (dmax - dmin) * (dmax - dmin);
            T sig, dsig = T(0);
            if (spread <= K3a) sig = T(1);
            else if (spread <= T(2) * K3a) {
                const T y = (spread / K3a - T(1));
                sig = T(2) * y * y * y - T(3) * y * y + T(1);
                dsig = (T(6) * y * y - T(6) * y) / K3a;   // d sig / d spread
            } else sig = T(0);
            T ph = T(1), dph = T(0), r = T(0);
            const bool use_max = dqg > tol;
            if (sig < T(1) && (dqg > tol || dqg < -tol)) {
                r = (use_max ? dmax : dmin) / dqg;
                ph = phi<T, PHI>(r, yt);
                dph = phi_deriv<T, PHI>(r, yt);
            }
            const T ll = llim[k];
            const T lsig = ll * (T(1) - ph);
            const T lr = ll * (T(1) - sig) * dph;
            T ldmax = T(0), ldmin = T(0), ldqg = T(0);
            if (dph != T(0)) {
                if (use_max) ldmax = lr / dqg; else ldmin = lr / dqg;
                ldqg = -lr * r / dqg;
            }
            const T lspread = lsig * dsig;
            ldmax = ldmax + T(2) * (dmax - dmin) * lspread;
            ldmin = ldmin - T(2) * (dmax - dmin) * lspread;
            lgx[k] = lgx[k] + ldqg * dx;
            lgy[k] = lgy[k] + ldqg * dy;
            if (GEO) {
                // dqg = gx dx + gy dy at the selected edge
                G2<T> a = ga.hc[hs];
                a.v[0] = a.v[0] + ldqg * gx[k];
                a.v[1] = a.v[1] + ldqg * gy[k];
                ga.hc[hs] = a;
                lcx = lcx - ldqg * gx[k];
                lcy = lcy - ldqg * gy[k];
                // sig(spread / K3a) with K3a = (K sqrt(A))^3: d sig / d K3a =
                // -dsig spread / K3a and d K3a / d A = 1.5 K3a / A
                lA = lA - lsig * dsig * spread / K3a * (T(1.5) * K3a / m.c_area[c]);
            }
            lqc[k] = lqc[k] - ldmax - ldmin;
            if (hmx[k] < 0) lqc[k] = lqc[k] + ldmax;
            if (hmn[k] < 0) lqc[k] = lqc[k] + ldmin;
#pragma unroll
            for (int j = 0; j < kMaxHe; ++j) {
                if (h0 + j == hmx[k]) nbc[j][k] = nbc[j][k] + ldmax;
                if (h0 + j == hmn[k]) nbc[j][k] = nbc[j][k] + ldmin;
            }
        }
    }
    // Green-Gauss gradient reversed (explicit.cpp:112-171)
    if (grad) {
        const T ia = m.c_inv_area[c];
        T sx[4], sy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) { sx[k] = lgx[k] * ia; sy[k] = lgy[k] * ia; }
        if (GEO) {
            // g = (1/A) S: d g / d A = -g / A
#pragma unroll
            for (int k = 0; k < 4; ++k) lA = lA - (lgx[k] * gx[k] + lgy[k] * gy[k]) / m.c_area[c];
        }
#pragma unroll
        for (int j = 0; j < kMaxHe; ++j) {
            const int h = h0 + j;
            if (h >= h1) continue;
            const int nb = m.he_nbr[h];
            if (nb == c) continue;
            const uint32_t ue = m.he_edge[h];
            const bool side = (ue & kSideBit) != 0;
            const Q4<T> geo = m.e_geo[ue & kEdgeMask];
            const T w = geo.v[3];
            const T w1 = T(1) - w;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                T lf = sx[k] * geo.v[0] + sy[k] * geo.v[1];
                if (side) lf = -lf;
                lf = lf * geo.v[2];
                // f = (qi w1 + qj w) len with qi the side-0 cell
                lqc[k] = lqc[k] + lf * (side ? w : w1);
                nbc[j][k] = nbc[j][k] + lf * (side ? w1 : w);
            }
            if (GEO) {
                // gx += +-f nx, gy += +-f ny with f = F len, F = qi (1-w) + qj w
                const Q4<T> qn = qin[nb];
                T lnx = T(0), lny = T(0), ll = T(0), lw = T(0);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const T qi = side ? qn.v[k] : qc.v[k];
                    const T qj = side ? qc.v[k] : qn.v[k];
                    const T F = qi * w1 + qj * w;
                    T lfk = sx[k] * geo.v[0] + sy[k] * geo.v[1];   // adjoint of f_k
                    T gxs = sx[k], gys = sy[k];
                    if (side) { lfk = -lfk; gxs = -gxs; gys = -gys; }
                    lnx = lnx + gxs * F * geo.v[2];
                    lny = lny + gys * F * geo.v[2];
                    ll = ll + lfk * F;
                    lw = lw + lfk * geo.v[2] * (qj - qi);
                }
                Q4<T> a = ga.he[h];
                a.v[0] = a.v[0] + lnx;
                a.v[1] = a.v[1] + lny;
                a.v[2] = a.v[2] + ll;
                a.v[3] = a.v[3] + lw;
                ga.he[h] = a;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kMaxHe; ++j)
        if (h0 + j < h1) nb_contrib[h0 + j] = Q4<T>{{nbc[j][0], nbc[j][1], nbc[j][2], nbc[j][3]}};
    Q4<T> o;
#pragma unroll
    for (int k = 0; k < 4; ++k) o.v[k] = lqc[k];
    lam_direct[c] = o;
    if (GEO) {
        Q4<T> a = ga.c[c];
        a.v[0] = a.v[0] + lA;
        a.v[1] = a.v[1] + lcx;
        a.v[2] = a.v[2] + lcy;
        ga.c[c] = a;
    }
}

// ---- gather pass ---------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_adj_gather(MeshView<T> m, const Q4<T>* __restrict__ lam_direct,
                                                       const Q4<T>* __restrict__ lam_face,
                                                       const Q4<T>* __restrict__ nb_contrib, Q4<T>* __restrict__ out,
                                                       int accumulate) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.n_total) return;
    T v[4] = {T(0), T(0), T(0), T(0)};
    const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
    if (c < m.n_real) {
        q4_add(v, lam_direct[c]);
    } else {
        // a ghost's face state is the ghost state
        for (int h = h0; h < h1; ++h) q4_add(v, lam_face[h]);
    }
    for (int h = h0; h < h1; ++h)
        if (m.he_nbr[h] < m.n_real) q4_add(v, nb_contrib[m.he_twin[h]]);
    Q4<T> o = accumulate ? out[c] : Q4<T>{{T(0), T(0), T(0), T(0)}};
    q4_add(o.v, Q4<T>{{v[0], v[1], v[2], v[3]}});
    out[c] = o;
}

// ---- time step reversed for one cell (the minimum of a global step) --------
// summed form: dt = cfl A / S, S = sum_h max(e_own, e_nbr) len_h
// edge-minimum form: dt = min(1, min_h cfl (A / len_h) / |e_h|)
template <class T, int KIND, bool GEO>
__global__ void k_adj_dt_cell(MeshView<T> m, PhysParams pp, const Q4<T>* __restrict__ q,
                              const int32_t* __restrict__ cell, const double* __restrict__ ld_blk, int n_blk,
                              Q4<T>* __restrict__ lam_q0, GeoAdj<T> ga) {
    if (blockIdx.x != 0) return;
    // sum of the per-tile partials, in a fixed pattern (deterministic)
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n_blk; i += kBlock) part[0] += ld_blk[i];
    block_sum4(part);
    const int c = cell[0];
    if (threadIdx.x != 0 || c < 0) return;
    const T lam_d = T(part[0]);
    if (KIND == 0 || lam_d == T(0)) return;
    const T g = T(pp.gamma);
    const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
    const Q4<T> qc = q[c];
    if (KIND == 1) {
        T S = T(0);
        for (int h = h0; h < h1; ++h) {
            const uint32_t ue = m.he_edge[h];
            const Q4<T> geo = m.e_geo[ue & kEdgeMask];
            const Q4<T> qn = q[m.he_nbr[h]];
            const bool side = (ue & kSideBit) != 0;
            const T eo = eig_sum(g, qc.v, geo.v[0], geo.v[1]);
            const T en = eig_sum(g, qn.v, geo.v[0], geo.v[1]);
            S = S + (side ? smax(en, eo) : smax(eo, en)) * geo.v[2];
        }
        const T dtc = T(pp.cfl) * m.c_area[c] / S;
        const T Sb = -lam_d * dtc / S;
        if (GEO) {
            Q4<T> a = ga.c[c];
            a.v[0] = a.v[0] + lam_d * dtc / m.c_area[c];
            ga.c[c] = a;
        }
        for (int h = h0; h < h1; ++h) {
            const uint32_t ue = m.he_edge[h];
            const Q4<T> geo = m.e_geo[ue & kEdgeMask];
            const int nb = m.he_nbr[h];
            const Q4<T> qn = q[nb];
            const bool side = (ue & kSideBit) != 0;
            const T eo = eig_sum(g, qc.v, geo.v[0], geo.v[1]);
            const T en = eig_sum(g, qn.v, geo.v[0], geo.v[1]);
            // which term the forward max kept: smax(a, b) = (a < b) ? b : a,
            // taken as smax(eo, en) on side 0 and smax(en, eo) on side 1
            const bool own = side ? (en < eo) : !(eo < en);
            const int who = own ? c : nb;
            T qb[4] = {T(0), T(0), T(0), T(0)};
            eig_sum_vjp(g, own ? qc.v : qn.v, geo.v[0], geo.v[1], Sb * geo.v[2], qb);
            Q4<T> l = lam_q0[who];
            for (int k = 0; k < 4; ++k) l.v[k] = l.v[k] + qb[k];
            lam_q0[who] = l;
            if (GEO) {
                // S = sum max(e_own, e_nbr) len: the length and, through e, the normal
                T nb2[2] = {T(0), T(0)};
                eig_sum_vjp_n(own ? qc.v : qn.v, geo.v[0], geo.v[1], Sb * geo.v[2], nb2);
                Q4<T> a = ga.he[h];
                a.v[0] = a.v[0] + nb2[0];
                a.v[1] = a.v[1] + nb2[1];
                a.v[2] = a.v[2] + Sb * (own ? eo : en);
                ga.he[h] = a;
            }
        }
    } else {
        T r = T(1);
        int hs = -1;
        for (int h = h0; h < h1; ++h) {
            const Q4<T> geo = m.e_geo[m.he_edge[h] & kEdgeMask];
            const T eo = eig_edge(g, qc.v, geo.v[0], geo.v[1]);
            const T val = T(pp.cfl) * (m.c_area[c] / geo.v[2]) / sabs(eo);
            if (val < r) { r = val; hs = h; }
        }
        if (hs < 0) return;
        const Q4<T> geo = m.e_geo[m.he_edge[hs] & kEdgeMask];
        const T eo = eig_edge(g, qc.v, geo.v[0], geo.v[1]);
        // val = K / |e|: d val / d e = -val sign(e) / |e|
        const T eb = -lam_d * r / sabs(eo) * ((eo < T(0)) ? T(-1) : T(1));
        T qb[4] = {T(0), T(0), T(0), T(0)};
        eig_edge_vjp(g, qc.v, geo.v[0], geo.v[1], eb, qb);
        Q4<T> l = lam_q0[c];
        for (int k = 0; k < 4; ++k) l.v[k] = l.v[k] + qb[k];
        lam_q0[c] = l;
        if (GEO) {
            // r = cfl (A / len) / |e|
            T nb2[2] = {T(0), T(0)};
            eig_edge_vjp_n(g, qc.v, eb, nb2);
            Q4<T> a = ga.he[hs];
            a.v[0] = a.v[0] + nb2[0];
            a.v[1] = a.v[1] + nb2[1];
            a.v[2] = a.v[2] - lam_d * r / geo.v[2];
            ga.he[hs] = a;
            Q4<T> b = ga.c[c];
            b.v[0] = b.v[0] + lam_d * r / m.c_area[c];
            ga.c[c] = b;
        }
    }
}

// ---- the cell whose time step is the global minimum ------------------------
// min_dt (explicit.cpp:555-564) keeps the first minimal entry in reference
// order; ref[d] is the reference index of device cell d.
__device__ inline bool better(double v, int32_t r, double bv, int32_t br) {
    return br < 0 || v < bv || (v == bv && r < br);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_argmin_blocks(const T* __restrict__ dt, const int32_t* __restrict__ ref,
                                                          int n, double* __restrict__ bv, int32_t* __restrict__ bref,
                                                          int32_t* __restrict__ bdev) {
    __shared__ double sv[kBlock];
    __shared__ int32_t sr[kBlock], sd[kBlock];
    double v = 0.0;
    int32_t r = -1, d = -1;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double x = (double)dt[i];
        if (x == x && better(x, ref[i], v, r)) { v = x; r = ref[i]; d = i; }
    }
    sv[threadIdx.x] = v; sr[threadIdx.x] = r; sd[threadIdx.x] = d;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const int o = threadIdx.x + off;
            if (sr[o] >= 0 && better(sv[o], sr[o], sv[threadIdx.x], sr[threadIdx.x])) {
                sv[threadIdx.x] = sv[o]; sr[threadIdx.x] = sr[o]; sd[threadIdx.x] = sd[o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { bv[blockIdx.x] = sv[0]; bref[blockIdx.x] = sr[0]; bdev[blockIdx.x] = sd[0]; }
}

__global__ __launch_bounds__(kBlock) void k_argmin_final(int nb, const double* __restrict__ bv,
                                                         const int32_t* __restrict__ bref,
                                                         const int32_t* __restrict__ bdev, int32_t* __restrict__ out) {
    __shared__ double sv[kBlock];
    __shared__ int32_t sr[kBlock], sd[kBlock];
    const int t = threadIdx.x;
    sv[t] = 0.0; sr[t] = -1; sd[t] = -1;
    if (t < nb) { sv[t] = bv[t]; sr[t] = bref[t]; sd[t] = bdev[t]; }
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (t < off) {
            const int o = t + off;
            if (sr[o] >= 0 && better(sv[o], sr[o], sv[t], sr[t])) { sv[t] = sv[o]; sr[t] = sr[o]; sd[t] = sd[o]; }
        }
        __syncthreads();
    }
    if (t == 0) out[0] = sd[0];
}

// ---- initial boundary update reversed (explicit.cpp:712) -------------------
template <class T, bool GEO>
__global__ __launch_bounds__(kBlock) void k_adj_bounds_init(MeshView<T> m, PhysParams pp,
                                                            const BcParam<T>* __restrict__ bcs,
                                                            const Q4<T>* __restrict__ q, Q4<T>* __restrict__ lam,
                                                            GeoAdj<T> ga) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.n_real) return;
    const T g = T(pp.gamma);
    const Q4<T> qc = q[c];
    T acc[4] = {T(0), T(0), T(0), T(0)};
    bool any = false;
    for (int h = m.he_off[c]; h < m.he_off[c + 1]; ++h) {
        const uint32_t ue = m.he_edge[h];
        if (!(ue & kBcBit)) continue;
        const Q4<T> geo = m.e_geo[ue & kEdgeMask];
        const Q4<T> lg = lam[m.he_nbr[h]];
        boundary_vjp(bcs[m.he_bc[h]], g, qc.v, geo.v[0], geo.v[1], lg.v, acc);
        if (GEO) {
            T nb2[2] = {T(0), T(0)};
            boundary_vjp_n(bcs[m.he_bc[h]], g, qc.v, geo.v[0], geo.v[1], lg.v, nb2);
            Q4<T> a = ga.he[h];
            a.v[0] = a.v[0] + nb2[0];
            a.v[1] = a.v[1] + nb2[1];
            ga.he[h] = a;
        }
        any = true;
    }
    if (any) {
        Q4<T> l = lam[c];
        q4_add(l.v, Q4<T>{{acc[0], acc[1], acc[2], acc[3]}});
        lam[c] = l;
    }
}

// ---- objective: J = sum_b p(ghost_b) (wx_b n_x + wy_b n_y) len_b -----------
// obj[b] = (ghost device cell, nx, ny, len*wx... ) packed as Q4: (nx, ny, len, unused) + weights
template <class T>
__global__ __launch_bounds__(kBlock) void k_objective(int n_b, PhysParams pp, const int32_t* __restrict__ ghost,
                                                      const Q4<T>* __restrict__ geo_w, const Q4<T>* __restrict__ q,
                                                      double* __restrict__ terms, Q4<T>* __restrict__ lam,
                                                      double* __restrict__ p_out) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= n_b) return;
    const Q4<T> gw = geo_w[b];   // (wx, wy, nx, ny) ... len folded into the weights
    const int gc = ghost[b];
    const Q4<T> qg = q[gc];
    const T g = T(pp.gamma);
    const T proj = gw.v[0] * gw.v[2] + gw.v[1] * gw.v[3];
    const T p = pressure(g, qg.v);
    terms[b] = (double)(p * proj);
    if (p_out) p_out[b] = (double)p;
    if (lam) {
        T l[4] = {T(0), T(0), T(0), T(0)};
        pressure_vjp(g, qg.v, proj, l);
        lam[gc] = Q4<T>{{l[0], l[1], l[2], l[3]}};
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_copy(const Q4<T>* __restrict__ a, Q4<T>* __restrict__ b, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = a[i];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_fill(Q4<T>* __restrict__ a, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = Q4<T>{{T(0), T(0), T(0), T(0)}};
}

}  // namespace

namespace {

// ---- smoothed global time step (FVH_FLAG_SMOOTH_KINKS) ------------------
// dt = (sum_i dt_i^-p)^(-1/p): d dt / d dt_i = (dt / dt_i)^(p+1).  Every
// cell's summed-CFL step (eig_sum_k, max_w) is reversed: into its own state
// (lam_q0, its own slot) and, per half-edge, into the neighbour's (nb_dt,
// gathered by k_adj_dt_gather).  lam_dt: the total adjoint of dt.
__global__ void k_adj_ld_sum(const double* __restrict__ ld_blk, int n_blk, double* __restrict__ out) {
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n_blk; i += kBlock) part[0] += ld_blk[i];
    block_sum4(part);
    if (threadIdx.x == 0) out[0] = part[0];
}

template <class T, bool GEO>
__global__ __launch_bounds__(kBlock) void k_adj_dt_soft(MeshView<T> m, PhysParams pp, const Q4<T>* __restrict__ q,
                                                        const T* __restrict__ dt, const T* __restrict__ dt_g,
                                                        const double* __restrict__ lam_dt,
                                                        Q4<T>* __restrict__ lam_q0, Q4<T>* __restrict__ nb_dt,
                                                        GeoAdj<T> ga) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.n_total) return;
    const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
    const T g = T(pp.gamma), wv = T(pp.kwv);
    T y = dt_g[0] / dt[c];
    T wgt = y;
    for (int i = 0; i < 7; ++i) wgt = wgt * wgt;   // y^128 (p = 128)
    const T lam_c = T(lam_dt[0]) * wgt * y;
    const Q4<T> qc = q[c];
    T S = T(0);
    for (int h = h0; h < h1; ++h) {
        const uint32_t ue = m.he_edge[h];
        const Q4<T> geo = m.e_geo[ue & kEdgeMask];
        const Q4<T> qn = q[m.he_nbr[h]];
        const bool side = (ue & kSideBit) != 0;
        const T eo = eig_sum_k(g, qc.v, geo.v[0], geo.v[1], wv);
        const T en = eig_sum_k(g, qn.v, geo.v[0], geo.v[1], wv);
        S = S + (side ? kmax(en, eo, wv) : kmax(eo, en, wv)) * geo.v[2];
    }
    // dt_c = cfl A / S
    const T dtc = T(pp.cfl) * m.c_area[c] / S;
    const T Sb = -lam_c * dtc / S;
    if (GEO) {
        Q4<T> a = ga.c[c];
        a.v[0] = a.v[0] + lam_c * dtc / m.c_area[c];
        ga.c[c] = a;
    }
    T own[4] = {T(0), T(0), T(0), T(0)};
    for (int h = h0; h < h1; ++h) {
        const uint32_t ue = m.he_edge[h];
        const Q4<T> geo = m.e_geo[ue & kEdgeMask];
        const Q4<T> qn = q[m.he_nbr[h]];
        const T eo = eig_sum_k(g, qc.v, geo.v[0], geo.v[1], wv);
        const T en = eig_sum_k(g, qn.v, geo.v[0], geo.v[1], wv);
        // max_w is symmetric: d/d eo = (1 + d)/2, d/d en = (1 - d)/2 with d = |.|_w'(eo - en)
        const T dk = kabs_d(eo - en, wv);
        const T eob = Sb * geo.v[2] * T(0.5) * (T(1) + dk);
        const T enb = Sb * geo.v[2] * T(0.5) * (T(1) - dk);
        eig_sum_vjp<T, true>(g, qc.v, geo.v[0], geo.v[1], eob, own, wv);
        T nq[4] = {T(0), T(0), T(0), T(0)};
        eig_sum_vjp<T, true>(g, qn.v, geo.v[0], geo.v[1], enb, nq, wv);
        nb_dt[h] = Q4<T>{{nq[0], nq[1], nq[2], nq[3]}};
        if (GEO) {
            // S = sum max_w(e_own, e_nbr) len: the length and, through e, the normal
            T nb2[2] = {T(0), T(0)};
            eig_sum_vjp_n<T, true>(qc.v, geo.v[0], geo.v[1], eob, nb2, wv);
            eig_sum_vjp_n<T, true>(qn.v, geo.v[0], geo.v[1], enb, nb2, wv);
            Q4<T> a = ga.he[h];
            a.v[0] = a.v[0] + nb2[0];
            a.v[1] = a.v[1] + nb2[1];
            a.v[2] = a.v[2] + Sb * kmax(eo, en, wv);
            ga.he[h] = a;
        }
    }
    Q4<T> l = lam_q0[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) l.v[k] = l.v[k] + own[k];
    lam_q0[c] = l;
}

// lam_q0 of each cell += the neighbours' time-step contributions to it, in
// its half-edge order (the neighbour's half-edge back to the cell is the twin)
template <class T>
__global__ __launch_bounds__(kBlock) void k_adj_dt_gather(MeshView<T> m, const Q4<T>* __restrict__ nb_dt,
                                                          Q4<T>* __restrict__ lam_q0) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.n_total) return;
    const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
    Q4<T> l = lam_q0[c];
    for (int h = h0; h < h1; ++h) {
        const Q4<T> a = nb_dt[m.he_twin[h]];
#pragma unroll
        for (int k = 0; k < 4; ++k) l.v[k] = l.v[k] + a.v[k];
    }
    lam_q0[c] = l;
}

}  // namespace

template <class T>
void launch_adj_dt_soft(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, const Q4<T>* q, const T* dt,
                        const T* dt_g, const double* ld_blk, int n_blk, double* lam_dt, Q4<T>* lam_q0, Q4<T>* nb_dt,
                        const GeoAdj<T>& ga) {
    if (m.n_total == 0) return;
    hipLaunchKernelGGL(k_adj_ld_sum, dim3(1), dim3(kBlock), 0, s, ld_blk, n_blk, lam_dt);
    const int nb = nblocks(m.n_total);
    if (ga.he)
        hipLaunchKernelGGL((k_adj_dt_soft<T, true>), dim3(nb), dim3(kBlock), 0, s, m, pp, q, dt, dt_g,
                           (const double*)lam_dt, lam_q0, nb_dt, ga);
    else
        hipLaunchKernelGGL((k_adj_dt_soft<T, false>), dim3(nb), dim3(kBlock), 0, s, m, pp, q, dt, dt_g,
                           (const double*)lam_dt, lam_q0, nb_dt, ga);
    hipLaunchKernelGGL((k_adj_dt_gather<T>), dim3(nb), dim3(kBlock), 0, s, m, (const Q4<T>*)nb_dt, lam_q0);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class T>
void launch_adj_stage(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const PhysParams& pp, int phi_kind,
                      bool grad, bool lim, bool lin, const BcParam<T>* bcs, int n_edges, const int4* epair,
                      const Q4<T>* qin, const Q4<T>* qout, const Q4<T>* gx, const Q4<T>* gy, const Q4<T>* lm,
                      const Q4<T>* qt, const T* dt, const T* dt_g, T alpha, const Q4<T>* lam_out, Q4<T>* lam_q0,
                      Q4<T>* lam_qt, Q4<T>* lam_bq, Q4<T>* lam_bincr, Q4<T>* lam_face, Q4<T>* lam_direct,
                      Q4<T>* nb_contrib, Q4<T>* lam_in, int accumulate_in, T* lam_dt, double* ld_blk,
                      const GeoAdj<T>& ga) {
    const int nrt = xcd_grid_blocks(tm.n_rt);
    if (nrt == 0) return;
    const bool geo = ga.he != nullptr;
    // the adjoint of this stage's output: explicit, or the parts the previous call left
    const LamSrc<T> L{lam_out, lam_direct, lam_face, nb_contrib};
    const int n_ghost = m.n_total - m.n_real;
    if (n_ghost > 0) {
#define FVH_AB(L_, G) hipLaunchKernelGGL((k_adj_bc<T, L_, G>), dim3(nblocks(n_ghost * kBcLanes)), dim3(kBlock), 0, s, m, pp, \
                                         bcs, qout, gx, gy, lm, L, lam_bq, lam_bincr, ga)
        if (lin) { if (geo) FVH_AB(true, true); else FVH_AB(true, false); }
        else { if (geo) FVH_AB(false, true); else FVH_AB(false, false); }
#undef FVH_AB
    }
    if (geo)
        hipLaunchKernelGGL((k_adj_update<T, true>), dim3(nrt), dim3(kBlock), 0, s, m, tm, pp, qt, dt, dt_g, alpha, L,
                           (const Q4<T>*)lam_bq, lam_q0, lam_qt, lam_dt, ld_blk, ga);
    else
        hipLaunchKernelGGL((k_adj_update<T, false>), dim3(nrt), dim3(kBlock), 0, s, m, tm, pp, qt, dt, dt_g, alpha, L,
                           (const Q4<T>*)lam_bq, lam_q0, lam_qt, lam_dt, ld_blk, ga);
    const int ne = nblocks(n_edges);
#define FVH_AF(L_, G, SM_) hipLaunchKernelGGL((k_adj_flux<T, L_, G, SM_>), dim3(ne), dim3(kBlock), 0, s, m, pp, \
                                              n_edges, epair, qin, gx, gy, lm, (const Q4<T>*)lam_qt, lam_face, ga)
    if (phi_kind == PHI_VENKAT_S) {
        // smoothed kinks: second order only
        if (geo) FVH_AF(true, true, true); else FVH_AF(true, false, true);
    } else if (lin) {
        if (geo) FVH_AF(true, true, false); else FVH_AF(true, false, false);
    } else {
        if (geo) FVH_AF(false, true, false); else FVH_AF(false, false, false);
    }
#undef FVH_AF
#define FVH_AG(P, L_, G) hipLaunchKernelGGL((k_adj_gradlim<T, P, L_, G>), dim3(xcd_grid(m.n_real)), dim3(kBlock), 0, s, \
                                            m, pp,                                                                   \
                                            (int)grad, (int)lim, qin, gx, gy, lm, (const Q4<T>*)lam_face,            \
                                            (const Q4<T>*)lam_bincr, lam_direct, nb_contrib, ga)
#define FVH_AGL(P)                                                                \
    do {                                                                          \
        if (lin) { if (geo) FVH_AG(P, true, true); else FVH_AG(P, true, false); } \
        else { if (geo) FVH_AG(P, false, true); else FVH_AG(P, false, false); }   \
    } while (0)
    switch (phi_kind) {
        case PHI_ZERO: FVH_AGL(PHI_ZERO); break;
        case PHI_MICHALAK: FVH_AGL(PHI_MICHALAK); break;
        case PHI_VENKAT: FVH_AGL(PHI_VENKAT); break;
        case PHI_VENKAT_S:
            if (geo) FVH_AG(PHI_VENKAT_S, true, true); else FVH_AG(PHI_VENKAT_S, true, false);
            break;
        default: FVH_AGL(PHI_MINMOD); break;
    }
#undef FVH_AGL
#undef FVH_AG
    // the adjoint of the stage input: gathered here only when asked for;
    // otherwise the next call gathers it on the fly
    if (lam_in)
        hipLaunchKernelGGL((k_adj_gather<T>), dim3(nblocks(m.n_total)), dim3(kBlock), 0, s, m,
                           (const Q4<T>*)lam_direct, (const Q4<T>*)lam_face, (const Q4<T>*)nb_contrib, lam_in,
                           accumulate_in);
}

template <class T>
void launch_dt_argmin(hipStream_t s, const T* dt, const int32_t* ref, int n, double* bv, int32_t* bref,
                      int32_t* bdev, int32_t* out) {
    const int nb = std::min(nblocks(n), kArgminBlocks);
    hipLaunchKernelGGL((k_argmin_blocks<T>), dim3(nb), dim3(kBlock), 0, s, dt, ref, n, bv, bref, bdev);
    hipLaunchKernelGGL(k_argmin_final, dim3(1), dim3(kBlock), 0, s, nb, (const double*)bv, (const int32_t*)bref,
                       (const int32_t*)bdev, out);
}

template <class T>
void launch_adj_dt(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, int dt_kind, const Q4<T>* q,
                   const int32_t* cell, const double* ld_blk, int n_blk, Q4<T>* lam_q0, const GeoAdj<T>& ga) {
    const bool geo = ga.he != nullptr;
#define FVH_AD(K, G) hipLaunchKernelGGL((k_adj_dt_cell<T, K, G>), dim3(1), dim3(kBlock), 0, s, m, pp, q, cell, ld_blk, \
                                        n_blk, lam_q0, ga)
    if (dt_kind == 1) { if (geo) FVH_AD(1, true); else FVH_AD(1, false); }
    else if (dt_kind == 2) { if (geo) FVH_AD(2, true); else FVH_AD(2, false); }
#undef FVH_AD
}

template <class T>
void launch_adj_bounds_init(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, const BcParam<T>* bcs,
                            const Q4<T>* q, Q4<T>* lam, const GeoAdj<T>& ga) {
    if (m.n_real == 0) return;
    if (ga.he)
        hipLaunchKernelGGL((k_adj_bounds_init<T, true>), dim3(nblocks(m.n_real)), dim3(kBlock), 0, s, m, pp, bcs, q,
                           lam, ga);
    else
        hipLaunchKernelGGL((k_adj_bounds_init<T, false>), dim3(nblocks(m.n_real)), dim3(kBlock), 0, s, m, pp, bcs, q,
                           lam, ga);
}

template <class T>
void launch_objective(hipStream_t s, int n_b, const PhysParams& pp, const int32_t* ghost, const Q4<T>* geo_w,
                      const Q4<T>* q, double* terms, Q4<T>* lam, double* p_out) {
    if (n_b == 0) return;
    hipLaunchKernelGGL((k_objective<T>), dim3(nblocks(n_b)), dim3(kBlock), 0, s, n_b, pp, ghost, geo_w, q, terms, lam,
                       p_out);
}

template <class T>
void launch_copy(hipStream_t s, const void* a, void* b, int n) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_copy<T>), dim3(nblocks(n)), dim3(kBlock), 0, s, (const Q4<T>*)a, (Q4<T>*)b, n);
}

template <class T>
void launch_fill(hipStream_t s, Q4<T>* a, int n) {
    if (n == 0) return;
    hipLaunchKernelGGL((k_fill<T>), dim3(nblocks(n)), dim3(kBlock), 0, s, a, n);
}

template void launch_adj_stage<double>(hipStream_t, const MeshView<double>&, const TileMap&, const PhysParams&, int,
                                       bool, bool, bool, const BcParam<double>*, int, const int4*,
                                       const Q4<double>*, const Q4<double>*, const Q4<double>*, const Q4<double>*,
                                       const Q4<double>*, const Q4<double>*, const double*, const double*, double,
                                       const Q4<double>*, Q4<double>*, Q4<double>*, Q4<double>*, Q4<double>*,
                                       Q4<double>*, Q4<double>*, Q4<double>*, Q4<double>*, int, double*, double*,
                                       const GeoAdj<double>&);
template void launch_adj_dt<double>(hipStream_t, const MeshView<double>&, const PhysParams&, int, const Q4<double>*,
                                    const int32_t*, const double*, int, Q4<double>*, const GeoAdj<double>&);
template void launch_adj_dt_soft<double>(hipStream_t, const MeshView<double>&, const PhysParams&, const Q4<double>*,
                                         const double*, const double*, const double*, int, double*, Q4<double>*,
                                         Q4<double>*, const GeoAdj<double>&);
template void launch_dt_argmin<double>(hipStream_t, const double*, const int32_t*, int, double*, int32_t*, int32_t*,
                                       int32_t*);
template void launch_adj_bounds_init<double>(hipStream_t, const MeshView<double>&, const PhysParams&,
                                             const BcParam<double>*, const Q4<double>*, Q4<double>*,
                                             const GeoAdj<double>&);
template void launch_objective<double>(hipStream_t, int, const PhysParams&, const int32_t*, const Q4<double>*,
                                       const Q4<double>*, double*, Q4<double>*, double*);
template void launch_fill<double>(hipStream_t, Q4<double>*, int);
template void launch_copy<double>(hipStream_t, const void*, void*, int);

}  // namespace fvh
