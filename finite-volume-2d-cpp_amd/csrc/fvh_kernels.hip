// fvh_kernels.hip -- gfx950 kernels of the explicit residual-and-update path.
// Built with -ffp-contract=off: no multiply-add is fused, so each FP64
// expression rounds exactly as the reference's x86-64 build does.
#include "fvh_device.h"

namespace fvh {

namespace {

// ---------------------------------------------------------------------------
// time step: calc_dt kinds + block minimum for the global-dt reduction
// (first-order configurations; second order computes dt in k_gradlim)
// ---------------------------------------------------------------------------
template <class T, int KIND, class S>
__device__ inline T cell_dt(const MeshView<T>& m, const PhysParams& pp, int c, const Q4<T>& qc,
                            const Q4<S>* __restrict__ q) {
    if (KIND == 0) return T(pp.dt_fixed);
    const T g = T(pp.gamma);
    const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
    if (KIND == 1 || KIND == 5) {
        // summed form: cfl*A / sum_e max(eig_i, eig_j)*len  (forwardStep/main.cpp:149-184);
        // KIND 5: its smoothed form, |.|_w and max_w (FVH_FLAG_SMOOTH_KINKS)
        const EigCell<T> eo_c = eig_cell(g, qc.v);
        const T wv = T(pp.kwv);
        T s = T(0);
        for (int hb = h0; hb < h1; hb += kHeChunk) {
            // a chunk of half-edges: records, geometry and neighbour states
            // loaded together (indices clamped to the cell's last half-edge)
            uint32_t uev[kHeChunk];
            int nbv[kHeChunk];
            Q4<T> geov[kHeChunk], qnv[kHeChunk];
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                const int hh = min(hb + i, h1 - 1);
                uev[i] = m.he_edge[hh];
                nbv[i] = m.he_nbr[hh];
            }
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                geov[i] = m.e_geo[uev[i] & kEdgeMask];
                qnv[i] = ldq<T>(q + nbv[i]);
            }
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                if (hb + i >= h1) break;
                const Q4<T>& geo = geov[i];
                const bool side = (uev[i] & kSideBit) != 0;
                if (KIND == 5) {
                    const T eo = eig_sum_k(eo_c, geo.v[0], geo.v[1], wv);
                    const T en = eig_sum_k(g, qnv[i].v, geo.v[0], geo.v[1], wv);
                    s = s + (side ? kmax(en, eo, wv) : kmax(eo, en, wv)) * geo.v[2];
                } else {
                    const T eo = eig_sum(eo_c, geo.v[0], geo.v[1]);
                    const T en = eig_sum(g, qnv[i].v, geo.v[0], geo.v[1]);
                    s = s + (side ? smax(en, eo) : smax(eo, en)) * geo.v[2];
                }
            }
        }
        return T(pp.cfl) * m.c_area[c] / s;
    }
    if (KIND == 4) {
        // summed viscous form: cfl*A / sum_e max(eig_i + 2 eigv_i, eig_j + 2 eigv_j)*len
        // (nacaTransient/main.cpp:243-285)
        const T mu = T(pp.mu), pr = T(pp.pr);
        T s = T(0);
        for (int h = h0; h < h1; ++h) {
            const uint32_t ue = m.he_edge[h];
            const Q4<T> geo = m.e_geo[ue & kEdgeMask];
            const int nb = m.he_nbr[h];
            const Q4<T> qn = ldq<T>(q + nb);
            const bool side = (ue & kSideBit) != 0;
            const T eo = eig_sum(g, qc.v, geo.v[0], geo.v[1]) + T(2) * eig_visc(g, mu, pr, qc.v, geo.v[2], m.c_area[c]);
            const T en = eig_sum(g, qn.v, geo.v[0], geo.v[1]) + T(2) * eig_visc(g, mu, pr, qn.v, geo.v[2], m.c_area[nb]);
            s = s + (side ? smax(en, eo) : smax(eo, en)) * geo.v[2];
        }
        return T(pp.cfl) * m.c_area[c] / s;
    }
    if (KIND == 3) {
        // edge-minimum viscous local form (flatPlate/main.cpp:239-277)
        const T mu = T(pp.mu), pr = T(pp.pr);
        const T A = m.c_area[c];
        T r = T(1);
        for (int h = h0; h < h1; ++h) {
            const Q4<T> geo = m.e_geo[m.he_edge[h] & kEdgeMask];
            const T ec = eig_abs(g, qc.v, geo.v[0], geo.v[1]) * geo.v[2];
            const T ev = eig_visc(g, mu, pr, qc.v, geo.v[2], A);
            r = smin(r, T(pp.cfl) * A / (ec + T(2) * ev));
        }
        return r;
    }
    // edge-minimum local form (naca/main.cpp:151-182)
    T r = T(1);
    const T A = m.c_area[c];
    for (int h = h0; h < h1; ++h) {
        const Q4<T> geo = m.e_geo[m.he_edge[h] & kEdgeMask];
        const T eo = eig_edge(g, qc.v, geo.v[0], geo.v[1]);
        r = smin(r, T(pp.cfl) * (A / geo.v[2]) / sabs(eo));
    }
    return r;
}

template <class T, int KIND, class S>
__global__ __launch_bounds__(kBlock) void k_dt(MeshView<T> m, TileMap tm, PhysParams pp,
                                               const Q4<S>* __restrict__ q, T* __restrict__ dt,
                                               double* __restrict__ blk_min, const Ctrl* __restrict__ ctrl) {
    const int lb = xcd_block(tm.n_all);
    if (lb < 0 || stopped(ctrl, tm.inst[lb])) return;
    const int c = tm.c0[lb] + threadIdx.x;
    double v = kInf;
    if (c < tm.c0[lb + 1]) {
        const T r = cell_dt<T, KIND>(m, pp, c, ldq<T>(q + c), q);
        dt[c] = r;
        v = (double)r;
    }
    v = block_min(v);
    if (threadIdx.x == 0) blk_min[lb] = v;
}

// One workgroup per instance: minimum over its real and ghost tiles.
template <class T>
__global__ __launch_bounds__(kBlock) void k_dt_min(InstView iv, const double* __restrict__ blk_min,
                                                   const T* __restrict__ dt_cells, T* __restrict__ dt_global,
                                                   const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x;
    if (stopped(ctrl, i)) return;
    const int4 r = iv.tiles[i];
    // four loads in flight per lane; the minimum of the same values in any
    // order (a running minimum from +inf never becomes NaN under smin)
    double v0 = kInf, v1 = kInf, v2 = kInf, v3 = kInf;
    int t = r.x + threadIdx.x;
    for (; t + 3 * kBlock < r.y; t += 4 * kBlock) {
        const double a = blk_min[t], b = blk_min[t + kBlock], c = blk_min[t + 2 * kBlock], d = blk_min[t + 3 * kBlock];
        v0 = smin(v0, a);
        v1 = smin(v1, b);
        v2 = smin(v2, c);
        v3 = smin(v3, d);
    }
    for (; t < r.y; t += kBlock) v0 = smin(v0, blk_min[t]);
    for (t = r.z + threadIdx.x; t < r.w; t += kBlock) v1 = smin(v1, blk_min[t]);
    double v = smin(smin(v0, v1), smin(v2, v3));
    v = block_min(v);
    if (threadIdx.x == 0) {
        // min_dt starts from dt[0] (explicit.cpp:557): a NaN there survives
        const T d0 = dt_cells[iv.dev0[i]];
        dt_global[i] = (d0 != d0) ? d0 : T(v);
    }
}

// Smoothed global time step (FVH_FLAG_SMOOTH_KINKS): the p-norm
// dt = (sum_i dt_i^-p)^(-1/p), p = 2^kSoftPow2, over an instance's cells
// (ghosts included, as min_dt), evaluated as m / S^(1/p) with m the minimum
// and S = sum_i (m / dt_i)^p.  Each term lies in [0, 1] and is added in
// 64.64 fixed point, so S does not depend on the summation order; S^(1/p) is
// kSoftPow2 square roots.  A NaN term (a NaN time step anywhere) makes the
// step NaN.
constexpr int kSoftPow2 = 7;
__device__ inline unsigned __int128 soft_term(double m, double d) {
    double y = m / d;
    if (!(y >= 0.0 && y <= 1.0)) return (unsigned __int128)1 << 100;
    for (int i = 0; i < kSoftPow2; ++i) y = y * y;
    if (y >= 1.0) return (unsigned __int128)1 << 64;
    return (unsigned __int128)(uint64_t)(y * 18446744073709551616.0);
}

__device__ inline unsigned __int128 block_sum_u128(unsigned __int128 v) {
    __shared__ unsigned __int128 red[kBlock];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    return red[0];
}

// per tile: the fixed-point sum of its cells' terms (two words per tile)
__global__ __launch_bounds__(kBlock) void k_dt_soft_tile(TileMap tm, const double* __restrict__ dt,
                                                         const double* __restrict__ dt_min,
                                                         uint64_t* __restrict__ part, const Ctrl* __restrict__ ctrl) {
    const int lb = xcd_block(tm.n_all);
    if (lb < 0 || stopped(ctrl, tm.inst[lb])) return;
    const int c = tm.c0[lb] + threadIdx.x;
    const unsigned __int128 v = c < tm.c0[lb + 1] ? soft_term(dt_min[tm.inst[lb]], dt[c]) : 0;
    const unsigned __int128 t = block_sum_u128(v);
    if (threadIdx.x == 0) {
        part[2 * lb] = (uint64_t)t;
        part[2 * lb + 1] = (uint64_t)(t >> 64);
    }
}

// per instance: S over its tiles, then dt = m / S^(1/p) in place of m
__global__ __launch_bounds__(kBlock) void k_dt_soft_inst(InstView iv, const uint64_t* __restrict__ part,
                                                         double* __restrict__ dt_global,
                                                         const Ctrl* __restrict__ ctrl) {
    const int i = blockIdx.x;
    if (stopped(ctrl, i)) return;
    const int4 r = iv.tiles[i];
    unsigned __int128 v = 0;
    for (int t = r.x + threadIdx.x; t < r.y; t += kBlock)
        v += ((unsigned __int128)part[2 * t + 1] << 64) | part[2 * t];
    for (int t = r.z + threadIdx.x; t < r.w; t += kBlock)
        v += ((unsigned __int128)part[2 * t + 1] << 64) | part[2 * t];
    const unsigned __int128 S = block_sum_u128(v);
    if (threadIdx.x == 0) {
        double root = (double)(uint64_t)(S >> 64) + (double)(uint64_t)S * 0x1p-64;
        for (int k = 0; k < kSoftPow2; ++k) root = sqrt(root);
        const double m = dt_global[i];
        dt_global[i] = (S >> 100) != 0 ? __builtin_nan("") : m / root;
    }
}

// Limiter of one real cell from its neighbour min/max and its gradient
// (explicit.cpp:207-257): lim[k] = min over the cell's half-edges of
// sigma + (1 - sigma) phi(r), starting from the reset value in lim.
// PHI_VENKAT_S (smoothed kinks): r's numerator blends dmax and dmin by
// t = dqg / sqrt(dqg^2 + w^2) instead of choosing by the sign of dqg,
// phi = venkat_dq regularised by w^2, sigma is sigma_smooth, and the running
// minimum is min_w.
template <class T, int PHI>
__device__ inline void limit_cell(const MeshView<T>& m, const PhysParams& pp, int c, int h0, int h1, const Q4<T>& qc,
                                  const G2<T>& cc, const T (&qmn)[4], const T (&qmx)[4], const T (&gx)[4],
                                  const T (&gy)[4], T (&lim)[4]) {
    const T Ka = T(pp.K) * m.c_sqrt_area[c];
    const T K3a = Ka * Ka * Ka;
    const T yt = T(pp.phi_yt);
    const T tol = T(1e-15);
    // the smoothing sigma depends on the cell and the variable only
    // (explicit.cpp:226-238): once per variable, the same expressions
    T sigs[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const T dmax = qmx[k] - qc.v[k];
        const T dmin = qmn[k] - qc.v[k];
        const T spread = (dmax - dmin) * (dmax - dmin);
        if (spread <= K3a) {
            sigs[k] = T(1);
        } else if (spread <= T(2) * K3a) {
            const T y = (spread / K3a - T(1));
            sigs[k] = PHI == PHI_VENKAT_S ? sigma_smooth(y) : sigma_cubic(y);
        } else {
            sigs[k] = T(0);
        }
    }
    for (int hb = h0; hb < h1; hb += kHeChunk) {
        // edge centres of a chunk of half-edges loaded together (indices
        // clamped to the cell's last half-edge)
        G2<T> ecv[kHeChunk];
#pragma unroll
        for (int i = 0; i < kHeChunk; ++i) ecv[i] = m.e_c[m.he_edge[min(hb + i, h1 - 1)] & kEdgeMask];
#pragma unroll
        for (int i = 0; i < kHeChunk; ++i) {
            if (hb + i >= h1) break;
            const G2<T> ec = ecv[i];
            const T dx = ec.v[0] - cc.v[0];
            const T dy = ec.v[1] - cc.v[1];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const T dqg = gx[k] * dx + gy[k] * dy;
                const T dmax = qmx[k] - qc.v[k];
                const T dmin = qmn[k] - qc.v[k];
                const T sig = sigs[k];
                T l = T(1);
                if (sig < T(1)) {
                    if (PHI == PHI_VENKAT_S) {
                        // smoothed kinks: the extremum by a blend on the sign of
                        // dqg, phi(d / dqg) without the quotient
                        const T w = T(pp.kw[k]);
                        const T t = dqg / sqrt(dqg * dqg + w * w);
                        const T d = T(0.5) * (dmax + dmin) + T(0.5) * (dmax - dmin) * t;
                        l = venkat_dq(d, dqg, w * w);
                    } else if (dqg > tol || dqg < -tol) {
                        // phi(dmax / dqg) or phi(dmin / dqg) by the sign of
                        // dqg: one quotient and one phi per lane, so a wave
                        // whose lanes differ in sign issues them once
                        const T num = dqg > tol ? dmax : dmin;
                        l = phi<T, PHI>(num / dqg, yt);
                    }
                }
                l = sig + (T(1) - sig) * l;
                lim[k] = PHI == PHI_VENKAT_S ? kmin(lim[k], l, T(pp.kwl)) : smin(lim[k], l);
            }
        }
    }
}

// The first chunk of a cell's half-edges as grad_lim_cell reads it (edge
// records, where the neighbours are read from, edge geometry), for callers
// that load it ahead.
template <class T>
struct HeFirst {
    uint32_t ue[kHeChunk];
    int nb[kHeChunk];
    Q4<T> geo[kHeChunk];
};

// ---------------------------------------------------------------------------
// Green-Gauss gradient + neighbour min/max + limiter of one real cell
// (explicit.cpp:112-258), optionally its time step.  nidx(h) yields where
// the neighbour across half-edge h is read from, nget(i, self) its state
// and whether it is the cell itself; the two-pass and the fused forms differ
// only in where they read from.
// ---------------------------------------------------------------------------
template <class T, int PHI, int DT, bool PRE = false, class NIdx, class NGet>
__device__ inline void grad_lim_cell(const MeshView<T>& m, const PhysParams& pp, int grad, int lim_on, bool real,
                                     bool do_dt, int c, int h0, int h1, const Q4<T>& qc, const G2<T>& cc, NIdx nidx,
                                     NGet nget, T (&gx)[4], T (&gy)[4], T (&lim)[4], T& dt_out,
                                     const HeFirst<T>* first = nullptr) {
    const T g = T(pp.gamma);
#pragma unroll
    for (int k = 0; k < 4; ++k) { gx[k] = T(0); gy[k] = T(0); }
    // limiters are reset to 1 when computed (explicit.cpp:184-186) and
    // stay zero-initialised otherwise (explicit.cpp:673)
    const T lim0 = lim_on ? T(1) : T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) lim[k] = lim0;
    T qmn[4], qmx[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) { qmn[k] = qc.v[k]; qmx[k] = qc.v[k]; }
    T dts = (DT == 2 || DT == 3) ? T(1) : T(0);
    // the cell's own wave-speed terms of the summed time step, once
    EigCell<T> eo_c{T(0), T(0), T(0)};
    if (DT == 1 && do_dt) eo_c = eig_cell(g, qc.v);
    // one pass over the half-edges feeds the Green-Gauss sums (edge order,
    // explicit.cpp:125-155), the neighbour min/max (:193-204) and the
    // time-step sum or minimum
    if ((real && (grad || lim_on)) || (DT && do_dt)) {
        // one chunk of half-edges from their records, neighbour slots and
        // edge geometry
        auto chunk = [&](int hb, const uint32_t(&uev)[kHeChunk], const int(&nbv)[kHeChunk],
                         const Q4<T>(&geov)[kHeChunk]) {
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                const int h = hb + i;
                if (h >= h1) break;
                const uint32_t ue = uev[i];
                const Q4<T> geo = geov[i];
                const bool side = (ue & kSideBit) != 0;
                bool self;
                const Q4<T> qn = nget(nbv[i], self);
                if (DT == 1 && do_dt) {
                    const T eo = eig_sum(eo_c, geo.v[0], geo.v[1]);
                    const T en = eig_sum(g, qn.v, geo.v[0], geo.v[1]);
                    dts = dts + (side ? smax(en, eo) : smax(eo, en)) * geo.v[2];
                } else if (DT == 2 && do_dt) {
                    const T eo = eig_edge(g, qc.v, geo.v[0], geo.v[1]);
                    dts = smin(dts, T(pp.cfl) * (m.c_area[c] / geo.v[2]) / sabs(eo));
                } else if (DT == 3 && do_dt) {
                    const T A = m.c_area[c];
                    const T ec = eig_abs(g, qc.v, geo.v[0], geo.v[1]) * geo.v[2];
                    const T ev = eig_visc(g, T(pp.mu), T(pp.pr), qc.v, geo.v[2], A);
                    dts = smin(dts, T(pp.cfl) * A / (ec + T(2) * ev));
                } else if (DT == 4 && do_dt) {
                    const T mu = T(pp.mu), pr = T(pp.pr);
                    const T eo = eig_sum(g, qc.v, geo.v[0], geo.v[1]) +
                                 T(2) * eig_visc(g, mu, pr, qc.v, geo.v[2], m.c_area[c]);
                    const T en = eig_sum(g, qn.v, geo.v[0], geo.v[1]) +
                                 T(2) * eig_visc(g, mu, pr, qn.v, geo.v[2], m.c_area[m.he_nbr[h]]);
                    dts = dts + (side ? smax(en, eo) : smax(eo, en)) * geo.v[2];
                }
                if (!real) continue;
                if (lim_on) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        if (PHI == PHI_VENKAT_S) {
                            qmn[k] = kmin(qmn[k], qn.v[k], T(pp.kw[k]));
                            qmx[k] = kmax(qmx[k], qn.v[k], T(pp.kw[k]));
                        } else {
                            qmn[k] = smin(qmn[k], qn.v[k]);
                            qmx[k] = smax(qmx[k], qn.v[k]);
                        }
                    }
                }
                if (!grad || self) continue;   // the reference skips i == j edges
                const T* qi = side ? qn.v : qc.v;
                const T* qj = side ? qc.v : qn.v;
                const T w = geo.v[3];
                const T w1 = T(1) - w;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const T f = (qi[k] * w1 + qj[k] * w) * geo.v[2];
                    if (!side) {
                        gx[k] = gx[k] + f * geo.v[0];
                        gy[k] = gy[k] + f * geo.v[1];
                    } else {
                        gx[k] = gx[k] - f * geo.v[0];
                        gy[k] = gy[k] - f * geo.v[1];
                    }
                }
            }
        };
        int hb0 = h0;
        if (PRE) {
            // the first chunk, loaded ahead by the caller (every cell has a
            // half-edge)
            chunk(h0, first->ue, first->nb, first->geo);
            hb0 += kHeChunk;
        }
        for (int hb = hb0; hb < h1; hb += kHeChunk) {
            // edge records and geometry of a chunk of half-edges loaded
            // together (indices clamped to the cell's last half-edge), so
            // that their latencies overlap
            uint32_t uev[kHeChunk];
            int nbv[kHeChunk];
            Q4<T> geov[kHeChunk];
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                const int hh = min(hb + i, h1 - 1);
                uev[i] = m.he_edge[hh];
                nbv[i] = nidx(hh);
            }
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) geov[i] = m.e_geo[uev[i] & kEdgeMask];
            chunk(hb, uev, nbv, geov);
        }
    }
    if (DT && do_dt) dt_out = (DT == 1 || DT == 4) ? T(pp.cfl) * m.c_area[c] / dts : dts;
    if (real && grad) {
        const T ia = m.c_inv_area[c];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gx[k] = gx[k] * ia;
            gy[k] = gy[k] * ia;
        }
    }
    if (real && lim_on) limit_cell<T, PHI>(m, pp, c, h0, h1, qc, cc, qmn, qmx, gx, gy, lim);
}

// Two-pass form, pass 1: gradient + limiter + face states of every cell of
// the tile, optionally its time step and the tile's minimum.
template <class T, int PHI, int DT, class S>
__global__ __launch_bounds__(kBlock) void k_gradlim(MeshView<T> m, TileMap tm, PhysParams pp, int grad, int lim_on,
                                                    int lin, const Q4<S>* __restrict__ q,
                                                    Q4<T>* __restrict__ face, Q4<T>* __restrict__ bincr,
                                                    T* __restrict__ dt, double* __restrict__ blk_min,
                                                    Q4<T>* __restrict__ dbg_gx, Q4<T>* __restrict__ dbg_gy,
                                                    Q4<T>* __restrict__ dbg_lim, const Ctrl* __restrict__ ctrl) {
    const int lb = xcd_block(tm.n_all);
    if (lb < 0 || stopped(ctrl, tm.inst[lb])) return;
    const int c = tm.c0[lb] + threadIdx.x;
    double vmin = kInf;
    if (c < tm.c0[lb + 1]) {
        const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
        const Q4<T> qc = ldq<T>(q + c);
        const G2<T> cc = m.c_c[c];
        const bool real = c < m.n_real;
        T gx[4], gy[4], lim[4], dtc = T(0);
        auto nidx = [&](int h) { return m.he_nbr[h]; };
        auto nget = [&](int nb, bool& self) {
            self = nb == c;
            return ldq<T>(q + nb);
        };
        grad_lim_cell<T, PHI, DT>(m, pp, grad, lim_on, real, true, c, h0, h1, qc, cc, nidx, nget, gx, gy, lim, dtc);
        if (DT) {
            dt[c] = dtc;
            vmin = (double)dtc;
        }
        if (dbg_gx) {
            Q4<T> a, b, d;
#pragma unroll
            for (int k = 0; k < 4; ++k) { a.v[k] = gx[k]; b.v[k] = gy[k]; d.v[k] = lim[k]; }
            dbg_gx[c] = a;
            dbg_gy[c] = b;
            if (dbg_lim) dbg_lim[c] = d;
        }
        if (lin) {
            for (int h = h0; h < h1; ++h) {
                const uint32_t ue = m.he_edge[h];
                const Q4<T> inc = face_inc(gx, gy, lim, m.e_c[ue & kEdgeMask], cc);
                Q4<T> fs;
#pragma unroll
                for (int k = 0; k < 4; ++k) fs.v[k] = qc.v[k] + inc.v[k];
                face[h] = fs;
                if (ue & kBcBit) bincr[m.he_nbr[h] - m.n_real] = inc;
            }
        }
    }
    if (DT) {
        vmin = block_min(vmin);
        if (threadIdx.x == 0) blk_min[lb] = vmin;
    }
}

// ---------------------------------------------------------------------------
// RK stage update + boundary update of one cell from its accumulated qt
// (explicit.cpp:368-417).  inc(h) is the limited increment of the cell's
// side of boundary half-edge h (this stage's gradient and limiter applied
// to the just-updated state: the reference's stale-gradient boundary).
// ---------------------------------------------------------------------------
template <class T, bool LIN, class S, class Inc>
__device__ __attribute__((always_inline)) inline void finish_cell(const MeshView<T>& m, const PhysParams& pp, const BcParam<T>* bcs, int c,
                                   int inst, int h0, int h1, const T (&qt)[4], const Q4<S>* __restrict__ q0,
                                   Q4<S>* __restrict__ qout, Inc inc_of,
                                   const T d, T alpha, bool want_rs,
                                   double (&rs)[4], Q4<T>* __restrict__ dbg_qt,
                                   const Q4<T>* qs_pre = nullptr) {
    const T g = T(pp.gamma);
    // the stage input, or the caller's copy of it loaded ahead
    const Q4<T> qs = qs_pre ? *qs_pre : ldq<T>(q0 + c);
    Q4<T> out;
#pragma unroll
    for (int k = 0; k < 4; ++k) out.v[k] = qs.v[k] + qt[k] * d * alpha;
    stq(qout + c, out);
    // rho > 0 and p > 0, i.e. 2 rho E > |rho u|^2 for rho > 0 (false for NaN):
    // a read-only check, the state is stored as computed
    if (pp.unphysical && !(out.v[0] > T(0) && T(2) * out.v[0] * out.v[3] > out.v[1] * out.v[1] + out.v[2] * out.v[2]))
        pp.unphysical[inst] = 1u;
    if (dbg_qt) {
        Q4<T> t;
#pragma unroll
        for (int k = 0; k < 4; ++k) t.v[k] = qt[k];
        dbg_qt[c] = t;
    }
    if (want_rs) {
        const double A = (double)m.c_area[c];
#pragma unroll
        for (int k = 0; k < 4; ++k) rs[k] = (double)qt[k] * (double)qt[k] * A;
    }
    // ghost cells of this cell's boundary edges
    for (int h = h0; h < h1; ++h) {
        const uint32_t ue = m.he_edge[h];
        if (!(ue & kBcBit)) continue;
        const int nb = m.he_nbr[h];
        const Q4<T> geo = m.e_geo[ue & kEdgeMask];
        T qi[4];
        if (LIN) {
            const Q4<T> ia = inc_of(h, nb);
#pragma unroll
            for (int k = 0; k < 4; ++k) qi[k] = out.v[k] + ia.v[k];
            if (pp.positivity && !physical(g, qi))
                for (int k = 0; k < 4; ++k) qi[k] = out.v[k];
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) qi[k] = out.v[k];
        }
        Q4<T> gs;
        boundary(bcs[m.he_bc[h]], g, gs.v, qi, geo.v[0], geo.v[1]);
        stq(qout + nb, gs);
    }
}

// Edge flux of kind FK (FVH_FLUX_*, or 3 and 4: the split and joined Roe
// fluxes with smoothed kinks) from the two face states; the viscous
// kind adds the laminar terms with the face gradient of cells ci (side 0)
// and cj (side 1), or a zero gradient without pp.diffusive
// (explicit.cpp:324-344).
template <class T, int FK, class S>
__device__ inline void edge_flux(const MeshView<T>& m, const PhysParams& pp, T g, T* f, const T* qi, const T* qj,
                                 T n0, T n1, int ci, int cj, const Q4<S>* __restrict__ qcell,
                                 const Q4<T>* __restrict__ gxc, const Q4<T>* __restrict__ gyc) {
    // FK 3, 4: the split and joined spellings with the smoothed entropy cure
    roe_flux<T, FK == 1 || FK == 4, (FK >= 3)>(g, f, qi, qj, n0, n1, T(pp.kwv));
    if (FK == 2) {
        T gxv[4] = {T(0), T(0), T(0), T(0)}, gyv[4] = {T(0), T(0), T(0), T(0)};
        if (pp.diffusive) {
            const Q4<T> gxi = ld(gxc + ci), gyi = ld(gyc + ci), gxj = ld(gxc + cj), gyj = ld(gyc + cj);
            const Q4<T> qa = ldq<T>(qcell + ci), qb = ldq<T>(qcell + cj);
            const G2<T> pa = m.c_c[ci], pb = m.c_c[cj];
            face_gradient(gxv, gyv, gxi.v, gyi.v, gxj.v, gyj.v, qa.v, qb.v, pa.v[0], pa.v[1], pb.v[0], pb.v[1]);
        }
        viscous_terms(g, T(pp.mu), T(pp.pr), T(pp.cp), T(pp.r), f, qi, qj, gxv, gyv, n0, n1);
    }
}

// Per-cell form: every interior edge flux is computed by both of its cells.
template <class T, int FK, bool LIN, class S>
__global__ __launch_bounds__(kBlock) void k_flux_cell(MeshView<T> m, TileMap tm, PhysParams pp,
                                                      const BcParam<T>* __restrict__ bcs,
                                                      const Q4<S>* __restrict__ qin, const Q4<S>* __restrict__ q0,
                                                      Q4<S>* __restrict__ qout, const Q4<T>* __restrict__ face,
                                                      const Q4<T>* __restrict__ bincr, const T* __restrict__ dt,
                                                      const T* __restrict__ dt_inst, T alpha,
                                                      double* __restrict__ res_blk,
                                                      Q4<T>* __restrict__ dbg_qt, const Ctrl* __restrict__ ctrl,
                                                      Q4<T>* __restrict__ qt_only, const Q4<T>* __restrict__ gxc,
                                                      const Q4<T>* __restrict__ gyc) {
    const int lb = xcd_block(tm.n_rt);
    if (lb < 0) return;
    const int inst = tm.inst[lb];
    if (stopped(ctrl, inst)) return;
    const int c = tm.c0[lb] + threadIdx.x;
    const T g = T(pp.gamma);
    double rs[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < tm.c0[lb + 1]) {
        const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
        const T A = m.c_area[c];
        T qt[4] = {T(0), T(0), T(0), T(0)};
        for (int h = h0; h < h1; ++h) {
            const uint32_t ue = m.he_edge[h];
            const Q4<T> geo = m.e_geo[ue & kEdgeMask];
            Q4<T> a, b;
            if (LIN) {
                a = ld(face + h);
                b = ld(face + m.he_twin[h]);
                if (pp.positivity) {
                    if (!physical(g, a.v)) a = ldq<T>(qin + c);
                    if (!physical(g, b.v)) b = ldq<T>(qin + m.he_nbr[h]);
                }
            } else {
                a = ldq<T>(qin + c);
                b = ldq<T>(qin + m.he_nbr[h]);
            }
            const bool side = (ue & kSideBit) != 0;
            T qi[4], qj[4], f[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                qi[k] = side ? b.v[k] : a.v[k];
                qj[k] = side ? a.v[k] : b.v[k];
            }
            const int nb = m.he_nbr[h];
            edge_flux<T, FK>(m, pp, g, f, qi, qj, geo.v[0], geo.v[1], side ? nb : c, side ? c : nb, qin, gxc, gyc);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const T t = f[k] * geo.v[2] / A;
                qt[k] = side ? qt[k] + t : qt[k] - t;
            }
        }
        if (qt_only) {
            // residual smoothing follows: hand qt over, the update comes later
            Q4<T> t;
#pragma unroll
            for (int k = 0; k < 4; ++k) t.v[k] = qt[k];
            qt_only[c] = t;
        } else {
            auto inc = [&](int, int nb) { return ld(bincr + (nb - m.n_real)); };
            finish_cell<T, LIN>(m, pp, bcs, c, inst, h0, h1, qt, q0, qout, inc, dt_inst ? dt_inst[inst] : dt[c], alpha,
                                res_blk != nullptr, rs, dbg_qt);
        }
    }
    if (res_blk) {
        block_sum4(rs);
        if (threadIdx.x == 0)
            for (int k = 0; k < 4; ++k) res_blk[4 * lb + k] = rs[k];
    }
}

// Tiled form: a workgroup owns a tile of 256 consecutive cells.  Phase 1
// computes the flux (times edge length) of every edge the tile touches once,
// edge-parallel, into LDS; phase 2 gathers each cell's edges from LDS in
// reference edge order.  Edges on a tile border are computed by both tiles
// (same inputs, same bits).
template <class T, int FK, bool LIN, class S>
__global__ __launch_bounds__(kBlock) void k_flux_tile(MeshView<T> m, TileMap tm, TileView tv, PhysParams pp,
                                                      const BcParam<T>* __restrict__ bcs,
                                                      const Q4<S>* __restrict__ qin, const Q4<S>* __restrict__ q0,
                                                      Q4<S>* __restrict__ qout, const Q4<T>* __restrict__ face,
                                                      const Q4<T>* __restrict__ bincr, const T* __restrict__ dt,
                                                      const T* __restrict__ dt_inst, T alpha,
                                                      double* __restrict__ res_blk,
                                                      Q4<T>* __restrict__ dbg_qt, const Ctrl* __restrict__ ctrl,
                                                      Q4<T>* __restrict__ qt_only, const Q4<T>* __restrict__ gxc,
                                                      const Q4<T>* __restrict__ gyc) {
    const int lb = xcd_block(tm.n_rt);
    if (lb < 0) return;
    const int inst = tm.inst[lb];
    if (stopped(ctrl, inst)) return;
    __shared__ Q4<T> fle[kTileEdges];
    const T g = T(pp.gamma);
    const int e0 = tv.off[lb], e1 = tv.off[lb + 1];
    for (int k = e0 + (int)threadIdx.x; k < e1; k += kBlock) {
        const Q4<T> geo = m.e_geo[tv.edge[k]];
        Q4<T> a, b;
        if (LIN) {
            const int2 hh = tv.he[k];
            a = ld(face + hh.x);
            b = ld(face + hh.y);
            if (pp.positivity) {
                if (!physical(g, a.v)) a = ldq<T>(qin + tv.cell[k].x);
                if (!physical(g, b.v)) b = ldq<T>(qin + tv.cell[k].y);
            }
        } else {
            const int2 cc = tv.cell[k];
            a = ldq<T>(qin + cc.x);
            b = ldq<T>(qin + cc.y);
        }
        T f[4];
        const int2 ce = tv.cell[k];
        edge_flux<T, FK>(m, pp, g, f, a.v, b.v, geo.v[0], geo.v[1], ce.x, ce.y, qin, gxc, gyc);
        Q4<T> r;
#pragma unroll
        for (int j = 0; j < 4; ++j) r.v[j] = f[j] * geo.v[2];
        fle[k - e0] = r;
    }
    __syncthreads();
    const int c = tm.c0[lb] + threadIdx.x;
    double rs[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < tm.c0[lb + 1]) {
        const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
        const T A = m.c_area[c];
        T qt[4] = {T(0), T(0), T(0), T(0)};
        for (int h = h0; h < h1; ++h) {
            const bool side = (m.he_edge[h] & kSideBit) != 0;
            const Q4<T> r = fle[m.he_slot[h]];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const T t = r.v[k] / A;
                qt[k] = side ? qt[k] + t : qt[k] - t;
            }
        }
        if (qt_only) {
            // residual smoothing follows: hand qt over, the update comes later
            Q4<T> t;
#pragma unroll
            for (int k = 0; k < 4; ++k) t.v[k] = qt[k];
            qt_only[c] = t;
        } else {
            auto inc = [&](int, int nb) { return ld(bincr + (nb - m.n_real)); };
            finish_cell<T, LIN>(m, pp, bcs, c, inst, h0, h1, qt, q0, qout, inc, dt_inst ? dt_inst[inst] : dt[c], alpha,
                                res_blk != nullptr, rs, dbg_qt);
        }
    }
    if (res_blk) {
        block_sum4(rs);
        if (threadIdx.x == 0)
            for (int k = 0; k < 4; ++k) res_blk[4 * lb + k] = rs[k];
    }
}

// One Jacobi sweep of the implicit residual smoothing of explicit.cpp:32-66,
//   qt_i = (qt0_i - s_i) / (1 + n_e eps),  s_i = sum over the cell's edges,
//   in edge order, of -(qt_j eps),  eps = 0.6,  n_e = 3 or 4 edges,
// where qt_j is the previous sweep's value (qt0 for the first) and a
// boundary ghost contributes its qt of 0 (explicit.cpp:352-362).  With
// qt_out the sweep's result is stored; without, it finishes the stage
// (update, boundary update, residual sums) as the flux pass would have.
template <class T, bool LIN, class S>
__global__ __launch_bounds__(kBlock) void k_smooth(MeshView<T> m, TileMap tm, PhysParams pp,
                                                   const BcParam<T>* __restrict__ bcs,
                                                   const Q4<T>* __restrict__ qt_nb, const Q4<T>* __restrict__ qt0,
                                                   Q4<T>* __restrict__ qt_out, const Q4<S>* __restrict__ q0,
                                                   Q4<S>* __restrict__ qout, const Q4<T>* __restrict__ bincr,
                                                   const T* __restrict__ dt, const T* __restrict__ dt_inst, T alpha,
                                                   double* __restrict__ res_blk, Q4<T>* __restrict__ dbg_qt,
                                                   const Ctrl* __restrict__ ctrl) {
    const int lb = xcd_block(tm.n_rt);
    if (lb < 0) return;
    const int inst = tm.inst[lb];
    if (stopped(ctrl, inst)) return;
    const int c = tm.c0[lb] + threadIdx.x;
    double rs[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < tm.c0[lb + 1]) {
        const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
        const T eps = T(0.6);
        T sm[4] = {T(0), T(0), T(0), T(0)};
        for (int h = h0; h < h1; ++h) {
            Q4<T> v;
            if (m.he_edge[h] & kBcBit) {
#pragma unroll
                for (int k = 0; k < 4; ++k) v.v[k] = T(0);
            } else {
                v = ld(qt_nb + m.he_nbr[h]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) sm[k] = sm[k] - v.v[k] * eps;
        }
        const T den = T(1) + T(h1 - h0) * eps;
        const Q4<T> a = ld(qt0 + c);
        T qt[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) qt[k] = (a.v[k] - sm[k]) / den;
        if (qt_out) {
            Q4<T> t;
#pragma unroll
            for (int k = 0; k < 4; ++k) t.v[k] = qt[k];
            qt_out[c] = t;
        } else {
            auto inc = [&](int, int nb) { return ld(bincr + (nb - m.n_real)); };
            finish_cell<T, LIN>(m, pp, bcs, c, inst, h0, h1, qt, q0, qout, inc, dt_inst ? dt_inst[inst] : dt[c], alpha,
                                res_blk != nullptr, rs, dbg_qt);
        }
    }
    if (res_blk) {
        block_sum4(rs);
        if (threadIdx.x == 0)
            for (int k = 0; k < 4; ++k) res_blk[4 * lb + k] = rs[k];
    }
}

// Fused form: one workgroup per tile does the whole stage.  Phase 0 stages
// the states of the tile and its two neighbour rings in LDS; phase 1
// computes gradient and limiter of the tile cells and of the real ring-1
// cells (a ring-1 cell's gradient is recomputed by every tile that borders
// it: same inputs, same order, same bits) and writes the reconstructed face
// state of each of their half-edges on a tile edge into that edge's slot;
// phase 2 evaluates each tile edge once from its two face states and leaves
// the flux in the slot; phase 3 gathers per cell in reference edge order and
// finishes the stage.  No face state goes through HBM; the boundary
// increments of the tile's own cells do (bincr), for phase 3.
// The stage of tile lb by one workgroup (LDS: the tile halo's states, then
// the edge slots).
template <class T, int PHI, int DT, bool JOINED, int NB, class S>
__device__ inline void stage_tile(int lb, unsigned char* lds_raw, const MeshView<T>& m, const TileMap& tm,
                                  const TileView& tv, const HaloView& hv, const PhysParams& pp, int grad, int lim_on,
                                  const BcParam<T>* __restrict__ bcs, const Q4<S>* __restrict__ qin,
                                  const Q4<S>* __restrict__ q0, Q4<S>* __restrict__ qout,
                                  Q4<T>* __restrict__ bincr, T* __restrict__ dt, const T* __restrict__ dt_inst,
                                  T alpha, double* __restrict__ res_blk, const Ctrl* __restrict__ ctrl,
                                  const StageRecord<T>& rec) {
    static_assert(NB >= kBlock, "phase 3 gives each tile cell one lane");
    const int inst = tm.inst[lb];
    if (stopped(ctrl, inst)) return;
    const int c0 = tm.c0[lb], nt = tm.c0[lb + 1] - c0;
    const int4 ex = hv.ext[lb];
    const int n_g = nt + ex.y, n_ext = nt + ex.z + ex.w;
    // states in the storage type, face states and fluxes in the arithmetic type
    Q4<S>* sq = (Q4<S>*)lds_raw;
    Q4<T>* fs = (Q4<T>*)(lds_raw + stage_lds_states<S>(n_ext));   // per tile edge slot: face state of side 0, side 1; then the flux in [0]
    const int32_t* xc = hv.cells + ex.x;
    const T g = T(pp.gamma);
    const T lim0 = lim_on ? T(1) : T(0);
    // Each phase starts with a chain of dependent global loads (cell, its
    // half-edge range, their records, the edge geometry).  The first links
    // of a lane's next item are issued before the barrier that ends the
    // previous phase, so that their latency overlaps the wait.
    struct CellHead {
        int c, h0, h1, r1;
    };
    auto cell_head = [&](int j) {
        CellHead x;
        const bool own = j < nt;
        x.c = own ? c0 + j : xc[j - nt];
        x.h0 = m.he_off[x.c];
        x.h1 = m.he_off[x.c + 1];
        x.r1 = own ? 0 : hv.r1_start[ex.x + j - nt];
        return x;
    };
    // a lane without a phase-1 cell takes the last one, so that every load
    // below stays in bounds
    const int j_first = min((int)threadIdx.x, n_g - 1);
    const CellHead head = cell_head(j_first);
    // phase 0: states of tile, ring 1 and ring 2.  A lane's first two
    // entries are loaded together (index loads, then state loads), so that
    // their latencies overlap; a lane without an entry loads the tile's
    // first cell (a tile has at least one) and stores nothing.
    {
        const int ja = threadIdx.x, jb = threadIdx.x + NB;
        const bool va = ja < n_ext, vb = jb < n_ext;
        const int ia = va ? (ja < nt ? c0 + ja : xc[ja - nt]) : c0;
        const int ib = vb ? (jb < nt ? c0 + jb : xc[jb - nt]) : c0;
        const Q4<S> sa = qin[ia], sb = qin[ib];
        if (va) sq[ja] = sa;
        if (vb) sq[jb] = sb;
        for (int j = threadIdx.x + 2 * NB; j < n_ext; j += NB) sq[j] = qin[j < nt ? c0 + j : xc[j - nt]];
    }
    // the next links of the lane's first phase-1 cell: its centre and the
    // first chunk of its half-edges (records, neighbour slots, geometry;
    // every cell has a half-edge)
    HeFirst<T> hf;
    {
        const uint16_t* loc = j_first < nt ? hv.he_loc + head.h0 : hv.r1_loc + head.r1;
#pragma unroll
        for (int i = 0; i < kHeChunk; ++i) {
            const int hh = min(head.h0 + i, head.h1 - 1);
            hf.ue[i] = m.he_edge[hh];
            hf.nb[i] = loc[hh - head.h0];
        }
#pragma unroll
        for (int i = 0; i < kHeChunk; ++i) hf.geo[i] = m.e_geo[hf.ue[i] & kEdgeMask];
    }
    const G2<T> cc_first = m.c_c[head.c];
    __syncthreads();
    // phase 1: gradients and limiters of the tile and its real ring-1 cells,
    // and their face states on the tile's edges; a lane's first cell from
    // the links loaded ahead (PRE), further ones (tiles whose cells with
    // their real ring 1 outnumber the lanes) from scratch
    const int e0 = tv.off[lb], e1 = tv.off[lb + 1];
    // positivity fallback on: each face state's pressure, computed for the
    // fallback test in phase 1, kept for the edge's flux in phase 2
    T* fpr = (T*)(fs + 2 * (e1 - e0));
    auto p1_cell = [&](int j, const CellHead& hd, auto pre) {
        constexpr bool PRE = decltype(pre)::value;
        const bool own = j < nt;
        const int c = hd.c, h0 = hd.h0, h1 = hd.h1, r1 = hd.r1;
        const uint16_t* loc = own ? hv.he_loc + h0 : hv.r1_loc + r1;
        T gx[4], gy[4], lim[4], dtc = T(0);
        auto nidx = [&](int h) { return (int)loc[h - h0]; };
        auto nget = [&](int l, bool& self) {
            self = l == j;
            return ldq<T>(sq + l);
        };
        const Q4<T> qc = ldq<T>(sq + j);
        const G2<T> cc = PRE ? cc_first : m.c_c[c];
        // the summed-CFL local step (DT 1) is formed after the face states
        // below, where the gradient phase's registers are free
        constexpr int DTG = DT == 1 ? 0 : DT;
        grad_lim_cell<T, PHI, DTG, PRE>(m, pp, grad, lim_on, true, own, c, h0, h1, qc, cc, nidx, nget, gx, gy, lim,
                                        dtc, &hf);
        if (DTG && own) dt[c] = dtc;
        if (rec.gx && own) {
            Q4<T> a, b, d;
#pragma unroll
            for (int k = 0; k < 4; ++k) { a.v[k] = gx[k]; b.v[k] = gy[k]; d.v[k] = lim[k]; }
            rec.gx[c] = a;
            rec.gy[c] = b;
            rec.lm[c] = d;
        }
        for (int hb = h0; hb < h1; hb += kHeChunk) {
            int slotv[kHeChunk];
            uint32_t uev[kHeChunk];
            G2<T> ecv[kHeChunk];
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                const int hh = min(hb + i, h1 - 1);
                slotv[i] = own ? (int)m.he_slot[hh] : (int)hv.r1_slot[r1 + (hh - h0)];
                uev[i] = m.he_edge[hh];
            }
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) ecv[i] = m.e_c[uev[i] & kEdgeMask];
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                const int h = hb + i;
                if (h >= h1) break;
                const int slot = slotv[i];
                const uint32_t ue = uev[i];
                const G2<T> ec = ecv[i];
                if (slot == 0xffff) continue;
                const int side = (ue & kSideBit) ? 1 : 0;
                const Q4<T> inc = face_inc(gx, gy, lim, ec, cc);
                Q4<T> f;
#pragma unroll
                for (int k = 0; k < 4; ++k) f.v[k] = qc.v[k] + inc.v[k];
                if (pp.positivity) {
                    // physical(): rho > 0 and p > 0
                    T pf = pressure(g, f.v);
                    if (!(f.v[0] > T(0) && pf > T(0))) {
                        f = qc;
                        pf = pressure(g, qc.v);
                    }
                    fpr[2 * slot + side] = pf;
                }
                fs[2 * slot + side] = f;
                if (own && (ue & kBcBit)) {
                    // the ghost's face: its state with a zero gradient and the reset limiter
                    const int nb = m.he_nbr[h];
                    bincr[nb - m.n_real] = inc;
                    const T z[4] = {T(0), T(0), T(0), T(0)};
                    const T l0[4] = {lim0, lim0, lim0, lim0};
                    if (rec.gx) {
                        // the ghost's gradient and limiter as the two-pass form stores them
                        rec.gx[nb] = Q4<T>{{z[0], z[1], z[2], z[3]}};
                        rec.gy[nb] = Q4<T>{{z[0], z[1], z[2], z[3]}};
                        rec.lm[nb] = Q4<T>{{l0[0], l0[1], l0[2], l0[3]}};
                    }
                    const Q4<T> ginc = face_inc(z, z, l0, ec, m.c_c[nb]);
                    const Q4<T> qg = ldq<T>(sq + loc[h - h0]);
                    Q4<T> fg;
#pragma unroll
                    for (int k = 0; k < 4; ++k) fg.v[k] = qg.v[k] + ginc.v[k];
                    if (pp.positivity) {
                        T pg = pressure(g, fg.v);
                        if (!(fg.v[0] > T(0) && pg > T(0))) {
                            fg = qg;
                            pg = pressure(g, qg.v);
                        }
                        fpr[2 * slot + 1 - side] = pg;
                    }
                    fs[2 * slot + 1 - side] = fg;
                }
            }
        }
        if (DT == 1 && own) {
            // summed-CFL local step (forwardStep/main.cpp:149-184): the
            // terms grad_lim_cell sums, in the same half-edge order
            const EigCell<T> eo_c = eig_cell(g, qc.v);
            T sum = T(0);
            for (int h = h0; h < h1; ++h) {
                const uint32_t ue = m.he_edge[h];
                const Q4<T> geo = m.e_geo[ue & kEdgeMask];
                const Q4<T> qn = ldq<T>(sq + loc[h - h0]);
                const bool side = (ue & kSideBit) != 0;
                const T eo = eig_sum(eo_c, geo.v[0], geo.v[1]);
                const T en = eig_sum(g, qn.v, geo.v[0], geo.v[1]);
                sum = sum + (side ? smax(en, eo) : smax(eo, en)) * geo.v[2];
            }
            dt[c] = T(pp.cfl) * m.c_area[c] / sum;
        }
    };
    if ((int)threadIdx.x < n_g) {
        p1_cell(threadIdx.x, head, std::integral_constant<bool, true>{});
        for (int j = threadIdx.x + NB; j < n_g; j += NB) p1_cell(j, cell_head(j), std::integral_constant<bool, false>{});
    }
    Q4<T> geo_next;
    if (e0 + (int)threadIdx.x < e1) geo_next = m.e_geo[tv.edge[e0 + threadIdx.x]];
    __syncthreads();
    // phase 2: each edge of the tile once, flux times length left in the slot
    for (int k = e0 + (int)threadIdx.x; k < e1; k += NB) {
        const Q4<T> geo = geo_next;
        if (k + NB < e1) geo_next = m.e_geo[tv.edge[k + NB]];
        const int sl = k - e0;
        const Q4<T> a = fs[2 * sl], b = fs[2 * sl + 1];
        T f[4];
        if (pp.positivity)
            roe_flux<T, JOINED, PHI == PHI_VENKAT_S, true>(g, f, a.v, b.v, geo.v[0], geo.v[1], T(pp.kwv),
                                                           fpr[2 * sl], fpr[2 * sl + 1]);
        else
            roe_flux<T, JOINED, PHI == PHI_VENKAT_S>(g, f, a.v, b.v, geo.v[0], geo.v[1], T(pp.kwv));
        Q4<T> r;
#pragma unroll
        for (int j = 0; j < 4; ++j) r.v[j] = f[j] * geo.v[2];
        fs[2 * sl] = r;
    }
    // phase 3's first links: the cell's stage input, time step (a local one
    // was stored in phase 1 by this same lane), half-edge range, records and
    // area
    const int j = threadIdx.x;
    int p3_h0 = 0, p3_h1 = 0;
    uint32_t p3_ue[kHeChunk];
    int p3_slot[kHeChunk];
    T p3_A = T(1), p3_dt = T(0);
    Q4<T> p3_q0;
    if (j < nt) {
        p3_q0 = ldq<T>(q0 + c0 + j);
        p3_dt = dt_inst ? dt_inst[inst] : dt[c0 + j];
        p3_h0 = m.he_off[c0 + j];
        p3_h1 = m.he_off[c0 + j + 1];
        p3_A = m.c_area[c0 + j];
        if (p3_h1 > p3_h0) {
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                const int hh = min(p3_h0 + i, p3_h1 - 1);
                p3_ue[i] = m.he_edge[hh];
                p3_slot[i] = m.he_slot[hh];
            }
        }
    }
    __syncthreads();
    // phase 3: ordered gather, RK update, boundary update
    double rs[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < nt) {
        const int c = c0 + j;
        const int h0 = p3_h0, h1 = p3_h1;
        const T A = p3_A;
        T qt[4] = {T(0), T(0), T(0), T(0)};
        for (int hb = h0; hb < h1; hb += kHeChunk) {
            uint32_t uev[kHeChunk];
            int slotv[kHeChunk];
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                const int hh = min(hb + i, h1 - 1);
                uev[i] = hb == h0 ? p3_ue[i] : m.he_edge[hh];
                slotv[i] = hb == h0 ? p3_slot[i] : (int)m.he_slot[hh];
            }
#pragma unroll
            for (int i = 0; i < kHeChunk; ++i) {
                if (hb + i >= h1) break;
                const bool side = (uev[i] & kSideBit) != 0;
                const Q4<T> r = fs[2 * slotv[i]];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const T t = r.v[k] / A;
                    qt[k] = side ? qt[k] + t : qt[k] - t;
                }
            }
        }
        auto inc = [&](int, int nb) { return ld(bincr + (nb - m.n_real)); };
        finish_cell<T, true>(m, pp, bcs, c, inst, h0, h1, qt, q0, qout, inc, p3_dt, alpha, res_blk != nullptr, rs,
                             rec.qt, &p3_q0);
    }
    if (res_blk) {
        block_sum4<NB>(rs);
        if (threadIdx.x == 0)
            for (int k = 0; k < 4; ++k) res_blk[4 * lb + k] = rs[k];
    }
}

// One launch grid over the tiles: workgroup b takes logical tile
// xcd_block(n_rt) (each XCD one contiguous range of tiles).  The stage that
// forms the summed-CFL local time step and the smoothed-kink stage ask for
// four waves per SIMD (at most 128 VGPRs; they took 131-137, three waves):
// the halos of the production meshes let four 256-lane workgroups share a
// CU, and they fit with a few registers spilled (measured faster, and
// bit-exact in the tests).  The other forms keep the compiler's choice.
template <class T, int PHI, int DT, bool JOINED, int NB, class S>
__global__ __launch_bounds__(NB, (DT == 1 || PHI == PHI_VENKAT_S) ? 4 : 1) void k_stage(MeshView<T> m, TileMap tm, TileView tv, HaloView hv, PhysParams pp,
                                                  int grad, int lim_on, const BcParam<T>* __restrict__ bcs,
                                                  const Q4<S>* __restrict__ qin, const Q4<S>* __restrict__ q0,
                                                  Q4<S>* __restrict__ qout, Q4<T>* __restrict__ bincr,
                                                  T* __restrict__ dt, const T* __restrict__ dt_inst, T alpha,
                                                  double* __restrict__ res_blk, const Ctrl* __restrict__ ctrl,
                                                  StageRecord<T> rec) {
    extern __shared__ __align__(32) unsigned char lds_raw[];
    const int lb = xcd_block(tm.n_rt);
    if (lb < 0) return;
    stage_tile<T, PHI, DT, JOINED, NB, S>(lb, lds_raw, m, tm, tv, hv, pp, grad, lim_on, bcs, qin, q0, qout, bincr, dt,
                                          dt_inst, alpha, res_blk, ctrl, rec);
}

// Persistent form: as many workgroups as the GPU holds at once, each pulling
// tiles from work queues until all are taken.  Queue x (x < kXcd) hands out
// the contiguous tile range the one-shot grid gives XCD x; workgroup b
// (resident on XCD b % kXcd) takes from its own XCD's queue, then from the
// others' in turn.  queue[0..kXcd-1] are ticket counters, queue[kXcd] counts
// the workgroups done; the last one resets all to zero, so the next launch
// (a graph replay too) starts from zero without a memset.  Thread 0 issues
// the atomic for a workgroup's next ticket before the current tile's work and
// reads it after, so its latency overlaps the tile.  Every workgroup leaves
// the loop once every queue is exhausted; there is no waiting on other
// workgroups.
template <class T, int PHI, int DT, bool JOINED, int NB, class S>
__global__ __launch_bounds__(NB) void k_stage_q(MeshView<T> m, TileMap tm, TileView tv, HaloView hv, PhysParams pp,
                                                    int grad, int lim_on, const BcParam<T>* __restrict__ bcs,
                                                    const Q4<S>* __restrict__ qin, const Q4<S>* __restrict__ q0,
                                                    Q4<S>* __restrict__ qout, Q4<T>* __restrict__ bincr,
                                                    T* __restrict__ dt, const T* __restrict__ dt_inst, T alpha,
                                                    double* __restrict__ res_blk, const Ctrl* __restrict__ ctrl,
                                                    StageRecord<T> rec, int32_t* __restrict__ queue) {
    extern __shared__ __align__(32) unsigned char lds_raw[];
    __shared__ int next_tile;
    const int n = tm.n_rt;
    const int per = (n + kXcd - 1) / kXcd;
    const int home = (int)(blockIdx.x % kXcd);
    const bool t0 = threadIdx.x == 0;
    // thread 0's ticket state: the queue in use, the queues found empty, the
    // pending ticket
    int qx = home, t_raw = 0;
    unsigned empty = 0;
    auto resolve = [&]() -> int {
        for (;;) {
            const int lo = qx * per, hi = min(lo + per, n);
            if (lo + t_raw < hi) return lo + t_raw;
            empty |= 1u << qx;
            int a = 1;
            while (a < kXcd && (empty & (1u << ((home + a) % kXcd)))) ++a;
            if (a == kXcd) return -1;
            qx = (home + a) % kXcd;
            t_raw = atomicAdd(&queue[qx], 1);
        }
    };
    int tile = -1;
    if (t0) {
        t_raw = atomicAdd(&queue[qx], 1);
        tile = resolve();
    }
    for (;;) {
        if (t0) {
            next_tile = tile;
            if (tile >= 0) t_raw = atomicAdd(&queue[qx], 1);   // the next ticket, read after this tile
        }
        __syncthreads();
        const int lb = next_tile;
        __syncthreads();
        if (lb < 0) break;
        stage_tile<T, PHI, DT, JOINED, NB, S>(lb, lds_raw, m, tm, tv, hv, pp, grad, lim_on, bcs, qin, q0, qout, bincr,
                                              dt, dt_inst, alpha, res_blk, ctrl, rec);
        if (t0) tile = resolve();
    }
    if (t0 && atomicAdd(&queue[kXcd], 1) == (int)gridDim.x - 1)
        for (int k = 0; k <= kXcd; ++k) atomicExch(&queue[k], 0);
}

// Boundary update from the cells' current states (update_bounds,
// explicit.cpp:380-417): with gradients and limiters given (gxa, gya, lima),
// or with zero ones when they are null, as the initial update of run() does
// (explicit.cpp:712).
template <class T, class S>
__global__ __launch_bounds__(kBlock) void k_bounds_init(MeshView<T> m, PhysParams pp, int lin,
                                                        const BcParam<T>* __restrict__ bcs, Q4<S>* __restrict__ q,
                                                        const Q4<T>* __restrict__ gxa, const Q4<T>* __restrict__ gya,
                                                        const Q4<T>* __restrict__ lima) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.n_real) return;
    const T g = T(pp.gamma);
    const Q4<T> qc = ldq<T>(q + c);
    const G2<T> cc = m.c_c[c];
    for (int h = m.he_off[c]; h < m.he_off[c + 1]; ++h) {
        const uint32_t ue = m.he_edge[h];
        if (!(ue & kBcBit)) continue;
        const int e = ue & kEdgeMask;
        const Q4<T> geo = m.e_geo[e];
        T qi[4];
        if (lin) {
            const G2<T> ec = m.e_c[e];
            const T dx = ec.v[0] - cc.v[0];
            const T dy = ec.v[1] - cc.v[1];
            const T zero = T(0);
            const Q4<T> gxc = gxa ? ld(gxa + c) : Q4<T>{{zero, zero, zero, zero}};
            const Q4<T> gyc = gya ? ld(gya + c) : Q4<T>{{zero, zero, zero, zero}};
            const Q4<T> lc = lima ? ld(lima + c) : Q4<T>{{zero, zero, zero, zero}};
#pragma unroll
            for (int k = 0; k < 4; ++k) qi[k] = qc.v[k] + (gxc.v[k] * dx + gyc.v[k] * dy) * lc.v[k];
            if (pp.positivity && !physical(g, qi))
                for (int k = 0; k < 4; ++k) qi[k] = qc.v[k];
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) qi[k] = qc.v[k];
        }
        Q4<T> gs;
        boundary(bcs[m.he_bc[h]], g, gs.v, qi, geo.v[0], geo.v[1]);
        stq(q + m.he_nbr[h], gs);
    }
}

// ---------------------------------------------------------------------------
// The phases one at a time on given arrays (the reference's per-phase
// functions), in device cell order.
// ---------------------------------------------------------------------------

// calc_limiters (explicit.cpp:174-258) from given gradients: qmin and qmax of
// every cell over its face neighbours, ghost cells included (:188-204), in
// the reference's edge order; the limiter of every real cell; ghost cells
// keep the reset value 1 (:184-186).
template <class T, int PHI>
__global__ __launch_bounds__(kBlock) void k_limiters(MeshView<T> m, PhysParams pp, const Q4<T>* __restrict__ q,
                                                     const Q4<T>* __restrict__ gxa, const Q4<T>* __restrict__ gya,
                                                     Q4<T>* __restrict__ lim_out, Q4<T>* __restrict__ qmin_out,
                                                     Q4<T>* __restrict__ qmax_out) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.n_total) return;
    const Q4<T> qc = ld(q + c);
    const int h0 = m.he_off[c], h1 = m.he_off[c + 1];
    T qmn[4], qmx[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) { qmn[k] = qc.v[k]; qmx[k] = qc.v[k]; }
    for (int h = h0; h < h1; ++h) {
        const Q4<T> qn = ld(q + m.he_nbr[h]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            qmn[k] = smin(qmn[k], qn.v[k]);
            qmx[k] = smax(qmx[k], qn.v[k]);
        }
    }
    T lim[4] = {T(1), T(1), T(1), T(1)};
    if (c < m.n_real) {
        const Q4<T> a = ld(gxa + c), b = ld(gya + c);
        T gx[4], gy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) { gx[k] = a.v[k]; gy[k] = b.v[k]; }
        limit_cell<T, PHI>(m, pp, c, h0, h1, qc, m.c_c[c], qmn, qmx, gx, gy, lim);
    }
    lim_out[c] = Q4<T>{{lim[0], lim[1], lim[2], lim[3]}};
    if (qmin_out) qmin_out[c] = Q4<T>{{qmn[0], qmn[1], qmn[2], qmn[3]}};
    if (qmax_out) qmax_out[c] = Q4<T>{{qmx[0], qmx[1], qmx[2], qmx[3]}};
}

// Reconstructed face state of every half-edge of every cell from given
// gradients and limiters, q + (gx dx + gy dy) lim (explicit.cpp:307-313),
// for the flux pass.
template <class T>
__global__ __launch_bounds__(kBlock) void k_faces(MeshView<T> m, const Q4<T>* __restrict__ q,
                                                  const Q4<T>* __restrict__ gxa, const Q4<T>* __restrict__ gya,
                                                  const Q4<T>* __restrict__ lima, Q4<T>* __restrict__ face) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.n_total) return;
    const Q4<T> qc = ld(q + c), a = ld(gxa + c), b = ld(gya + c), l = ld(lima + c);
    const G2<T> cc = m.c_c[c];
    T gx[4], gy[4], lim[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) { gx[k] = a.v[k]; gy[k] = b.v[k]; lim[k] = l.v[k]; }
    for (int h = m.he_off[c]; h < m.he_off[c + 1]; ++h) {
        const Q4<T> inc = face_inc(gx, gy, lim, m.e_c[m.he_edge[h] & kEdgeMask], cc);
        Q4<T> fs;
#pragma unroll
        for (int k = 0; k < 4; ++k) fs.v[k] = qc.v[k] + inc.v[k];
        face[h] = fs;
    }
}

// update_cells (explicit.cpp:368-378) over n values: q = ql + qt dt v.
__global__ __launch_bounds__(kBlock) void k_update_cells(double* __restrict__ q, const double* __restrict__ ql,
                                                         const double* __restrict__ qt,
                                                         const double* __restrict__ dt, double v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) q[i] = ql[i] + qt[i] * dt[i] * v;
}

// min_dt (explicit.cpp:555-564) over n values: per-block minima (a NaN never
// wins, as std::min(min, dt[i]) keeps min), then the minimum of the blocks,
// NaN if dt[0] is NaN (the reference starts from dt[0]), written to every
// value.
__global__ __launch_bounds__(kBlock) void k_array_min(const double* __restrict__ a, int64_t n,
                                                      double* __restrict__ part) {
    double v = kInf;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        v = smin(v, a[i]);
    v = block_min(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ __launch_bounds__(kBlock) void k_array_min_fill(double* __restrict__ a, int64_t n,
                                                           const double* __restrict__ part, int n_part) {
    __shared__ double r;
    double v = kInf;
    for (int i = threadIdx.x; i < n_part; i += kBlock) v = smin(v, part[i]);
    v = block_min(v);
    if (threadIdx.x == 0) r = (a[0] != a[0]) ? a[0] : v;
    __syncthreads();
    const double w = r;
    // every block reads r before any value changes: one block fills all
    for (int64_t i = threadIdx.x; i < n; i += kBlock) a[i] = w;
}

// calc_residuals (explicit.cpp:481-552) from a given qt: per real tile the
// sum of qt^2 A in a fixed order, then per instance the square root of the
// sum over its tiles.
template <class T>
__global__ __launch_bounds__(kBlock) void k_res_tile(MeshView<T> m, TileMap tm, const Q4<T>* __restrict__ qt,
                                                     double* __restrict__ blk) {
    const int t = blockIdx.x;
    if (t >= tm.n_rt) return;
    const int c = tm.c0[t] + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < tm.c0[t + 1]) {
        const Q4<T> a = ld(qt + c);
        const double A = (double)m.c_area[c];
        for (int k = 0; k < 4; ++k) v[k] = (double)a.v[k] * (double)a.v[k] * A;
    }
    block_sum4(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) blk[4 * t + k] = v[k];
}

__global__ __launch_bounds__(kBlock) void k_res_inst(InstView iv, const double* __restrict__ blk,
                                                     double* __restrict__ out) {
    const int i = blockIdx.x;
    const int4 r = iv.tiles[i];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = r.x + threadIdx.x; t < r.y; t += kBlock)
        for (int k = 0; k < 4; ++k) s[k] += blk[4 * t + k];
    block_sum4(s);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) out[4 * i + k] = sqrt(s[k]);
}

// ---------------------------------------------------------------------------
// run() bookkeeping on the device (explicit.cpp:715-801)
// ---------------------------------------------------------------------------
__global__ void k_ctrl_begin(int n_inst, Ctrl* ctrl, uint32_t max_step, double max_time, double tol, int stop_var) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    Ctrl* c = ctrl + i;
    if (c->stop) return;
    double Rmax = 0.0;
    if (c->step == 0) {
        Rmax = 1.0;
    } else if (stop_var >= 0 && stop_var < 4) {
        Rmax = c->R[stop_var];
    } else {
        for (int k = 0; k < 4; ++k) Rmax = smax(c->R[k], Rmax);
    }
    if ((c->step >= max_step) | (c->time >= max_time) | (Rmax < tol)) c->stop = 1;
}

// One workgroup per instance: residual norms from the per-tile partial sums
// (explicit.cpp:481-552, 752-771) and the step/time bookkeeping (:800).
template <class T>
__global__ __launch_bounds__(kBlock) void k_ctrl_end(InstView iv, Ctrl* ctrl, const double* __restrict__ res_blk,
                                                     int need_res, const T* __restrict__ dt_src, int dt_is_inst) {
    const int i = blockIdx.x;
    Ctrl* c = ctrl + i;
    if (c->stop) return;
    if (need_res) {
        const int4 r = iv.tiles[i];
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int t = r.x + threadIdx.x; t < r.y; t += kBlock)
            for (int k = 0; k < 4; ++k) s[k] += res_blk[4 * t + k];
        block_sum4(s);
        if (threadIdx.x == 0) {
            double R[4];
            for (int k = 0; k < 4; ++k) R[k] = sqrt(s[k]);
            if (c->step == 0) {
                for (int k = 0; k < 4; ++k) { c->R0[k] = R[k]; c->R[k] = R[k]; }
            } else {
                for (int k = 0; k < 4; ++k) c->R[k] = R[k] / c->R0[k];
            }
        }
    }
    if (threadIdx.x == 0) {
        const double d0 = (double)dt_src[dt_is_inst ? i : iv.dev0[i]];
        c->dt0 = d0;
        c->step += 1;
        c->time_prev = c->time;
        c->time += d0;
        for (int k = 0; k < 4; ++k)
            if (!(c->R[k] == c->R[k])) c->nonfinite = 1;
    }
}

// Conserved sums: per real tile the FP64 partial sum_c A_c q_ck in a fixed
// order, then one workgroup per instance adds its tiles' partials.
template <class T, class S>
__global__ __launch_bounds__(kBlock) void k_sums_tile(MeshView<T> m, TileMap tm, const Q4<S>* __restrict__ q,
                                                      double* __restrict__ blk) {
    const int t = blockIdx.x;
    if (t >= tm.n_rt) return;
    const int c = tm.c0[t] + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (c < tm.c0[t + 1]) {
        const Q4<S> qc = q[c];
        const double A = (double)m.c_area[c];
        for (int k = 0; k < 4; ++k) v[k] = A * (double)qc.v[k];
    }
    block_sum4(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) blk[4 * t + k] = v[k];
}

__global__ __launch_bounds__(kBlock) void k_sums_inst(InstView iv, const double* __restrict__ blk,
                                                      double* __restrict__ out) {
    const int i = blockIdx.x;
    const int4 r = iv.tiles[i];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = r.x + threadIdx.x; t < r.y; t += kBlock)
        for (int k = 0; k < 4; ++k) s[k] += blk[4 * t + k];
    block_sum4(s);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) out[4 * i + k] = s[k];
}


}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class T, class S>
void launch_dt(hipStream_t s, const MeshView<T>& m, const TileMap& tm, int dt_kind, const PhysParams& pp,
               const Q4<S>* q, T* dt, double* blk_min, const Ctrl* ctrl) {
    const int nb = xcd_grid_blocks(tm.n_all);
    if (nb == 0) return;
    if (dt_kind == 5) {
        if constexpr (kSmoothable<T, S>)
            hipLaunchKernelGGL((k_dt<T, 5, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, pp, q, dt, blk_min, ctrl);
        return;
    }
    if (dt_kind == 0) hipLaunchKernelGGL((k_dt<T, 0, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, pp, q, dt, blk_min, ctrl);
    else if (dt_kind == 1) hipLaunchKernelGGL((k_dt<T, 1, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, pp, q, dt, blk_min, ctrl);
    else if (dt_kind == 3) hipLaunchKernelGGL((k_dt<T, 3, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, pp, q, dt, blk_min, ctrl);
    else if (dt_kind == 4) hipLaunchKernelGGL((k_dt<T, 4, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, pp, q, dt, blk_min, ctrl);
    else hipLaunchKernelGGL((k_dt<T, 2, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, pp, q, dt, blk_min, ctrl);
}

void launch_dt_soft(hipStream_t s, const TileMap& tm, const InstView& iv, const double* dt_cells, double* dt_global,
                    uint64_t* part, const Ctrl* ctrl) {
    if (iv.n == 0 || tm.n_all == 0) return;
    hipLaunchKernelGGL(k_dt_soft_tile, dim3(xcd_grid_blocks(tm.n_all)), dim3(kBlock), 0, s, tm, dt_cells,
                       (const double*)dt_global, part, ctrl);
    hipLaunchKernelGGL(k_dt_soft_inst, dim3(iv.n), dim3(kBlock), 0, s, iv, (const uint64_t*)part, dt_global, ctrl);
}

void launch_dt_min(hipStream_t s, const InstView& iv, const double* blk_min, const void* dt_cells, int dt_is_f32,
                   void* dt_global, const Ctrl* ctrl) {
    if (iv.n == 0) return;
    if (dt_is_f32)
        hipLaunchKernelGGL((k_dt_min<float>), dim3(iv.n), dim3(kBlock), 0, s, iv, blk_min, (const float*)dt_cells,
                           (float*)dt_global, ctrl);
    else
        hipLaunchKernelGGL((k_dt_min<double>), dim3(iv.n), dim3(kBlock), 0, s, iv, blk_min, (const double*)dt_cells,
                           (double*)dt_global, ctrl);
}

template <class T, class S>
void launch_gradlim(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const PhysParams& pp, int phi_kind,
                    bool grad, bool lim, bool lin, const Q4<S>* q, Q4<T>* face, Q4<T>* bincr, int dt_kind, T* dt,
                    double* blk_min, Q4<T>* dbg_gx, Q4<T>* dbg_gy, Q4<T>* dbg_lim, const Ctrl* ctrl) {
    const int nb = xcd_grid_blocks(tm.n_all);
    if (nb == 0) return;
#define FVH_GL(P, D) hipLaunchKernelGGL((k_gradlim<T, P, D, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, pp, (int)grad, \
                                        (int)lim, (int)lin, q, face, bincr, dt, blk_min, dbg_gx, dbg_gy, dbg_lim, ctrl)
#define FVH_GLD(P)                                  \
    do {                                            \
        if (dt_kind == 1) FVH_GL(P, 1);             \
        else if (dt_kind == 2) FVH_GL(P, 2);        \
        else if (dt_kind == 3) FVH_GL(P, 3);        \
        else if (dt_kind == 4) FVH_GL(P, 4);        \
        else FVH_GL(P, 0);                          \
    } while (0)
    switch (phi_kind) {
        case PHI_ZERO: FVH_GLD(PHI_ZERO); break;
        case PHI_MICHALAK: FVH_GLD(PHI_MICHALAK); break;
        case PHI_VENKAT: FVH_GLD(PHI_VENKAT); break;
        case PHI_VENKAT_S:
            // smoothed kinks: FP64 only, the time step from the time-step kernel
            if constexpr (kSmoothable<T, S>) FVH_GL(PHI_VENKAT_S, 0);
            break;
        default: FVH_GLD(PHI_MINMOD); break;
    }
#undef FVH_GLD
#undef FVH_GL
}

template <class T, class S>
void launch_flux(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const TileView* tiles, const PhysParams& pp,
                 int flux_kind, bool lin, const BcParam<T>* bcs, const Q4<S>* qin, const Q4<S>* q0, Q4<S>* qout,
                 const Q4<T>* face, const Q4<T>* bincr, const T* dt, const T* dt_inst, T alpha, double* res_blk,
                 Q4<T>* dbg_qt, const Ctrl* ctrl, Q4<T>* qt_only, const Q4<T>* gxc, const Q4<T>* gyc) {
    const int nb = xcd_grid_blocks(tm.n_rt);
    if (nb == 0) return;
    if (tiles) {
        const TileView tv = *tiles;
#define FVH_FT(J, L) hipLaunchKernelGGL((k_flux_tile<T, J, L, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, tv, pp, bcs, \
                                        qin, q0, qout, face, bincr, dt, dt_inst, alpha, res_blk, dbg_qt, ctrl, \
                                        qt_only, gxc, gyc)
        if (flux_kind == 1) {
            if (lin) FVH_FT(1, true); else FVH_FT(1, false);
        } else if (flux_kind == 2) {
            if (lin) FVH_FT(2, true); else FVH_FT(2, false);
        } else if (flux_kind == 3 || flux_kind == 4) {
            if constexpr (kSmoothable<T, S>) {
                if (flux_kind == 3) { if (lin) FVH_FT(3, true); else FVH_FT(3, false); }
                else { if (lin) FVH_FT(4, true); else FVH_FT(4, false); }
            }
        } else {
            if (lin) FVH_FT(0, true); else FVH_FT(0, false);
        }
#undef FVH_FT
        return;
    }
#define FVH_FX(J, L) hipLaunchKernelGGL((k_flux_cell<T, J, L, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, pp, bcs, qin, \
                                        q0, qout, face, bincr, dt, dt_inst, alpha, res_blk, dbg_qt, ctrl, qt_only, \
                                        gxc, gyc)
    if (flux_kind == 1) {
        if (lin) FVH_FX(1, true); else FVH_FX(1, false);
    } else if (flux_kind == 2) {
        if (lin) FVH_FX(2, true); else FVH_FX(2, false);
    } else if (flux_kind == 3 || flux_kind == 4) {
        if constexpr (kSmoothable<T, S>) {
            if (flux_kind == 3) { if (lin) FVH_FX(3, true); else FVH_FX(3, false); }
            else { if (lin) FVH_FX(4, true); else FVH_FX(4, false); }
        }
    } else {
        if (lin) FVH_FX(0, true); else FVH_FX(0, false);
    }
#undef FVH_FX
}

template <class T, class S>
void launch_smooth(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const PhysParams& pp, bool lin,
                   const BcParam<T>* bcs, const Q4<T>* qt_nb, const Q4<T>* qt0, Q4<T>* qt_out, const Q4<S>* q0,
                   Q4<S>* qout, const Q4<T>* bincr, const T* dt, const T* dt_inst, T alpha, double* res_blk,
                   Q4<T>* dbg_qt, const Ctrl* ctrl) {
    const int nb = xcd_grid_blocks(tm.n_rt);
    if (nb == 0) return;
#define FVH_SM(L) hipLaunchKernelGGL((k_smooth<T, L, S>), dim3(nb), dim3(kBlock), 0, s, m, tm, pp, bcs, qt_nb, qt0, \
                                     qt_out, q0, qout, bincr, dt, dt_inst, alpha, res_blk, dbg_qt, ctrl)
    if (lin) FVH_SM(true); else FVH_SM(false);
#undef FVH_SM
}

template <class T, class S>
size_t stage_lds_bytes(const HaloView& hv) {
    return hv.lds_bytes > 0 ? (size_t)hv.lds_bytes : stage_lds_tile<T, S>(hv.max_ext, hv.max_edges, true);
}

template <class T, int PHI, int DT, bool J, int NB, class S>
void launch_stage_nb(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const TileView& tv, const HaloView& hv,
                     const PhysParams& pp, bool grad, bool lim, const BcParam<T>* bcs, const Q4<S>* qin,
                     const Q4<S>* q0, Q4<S>* qout, Q4<T>* bincr, T* dt, const T* dt_inst, T alpha,
                     double* res_blk, const Ctrl* ctrl, size_t lds, const StageRecord<T>& rec, int32_t* queue) {
    static bool attr_set = false;
    static int resident = 0;   // workgroups of the persistent form the GPU holds at once at this LDS size
    static size_t resident_lds = 0;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void*)k_stage<T, PHI, DT, J, NB, S>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                160 * 1024) != hipSuccess)
            (void)hipGetLastError();
        if (hipFuncSetAttribute((const void*)k_stage_q<T, PHI, DT, J, NB, S>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
            (void)hipGetLastError();
        attr_set = true;
    }
    if (!queue) {
        const int nb = xcd_grid_blocks(tm.n_rt);
        hipLaunchKernelGGL((k_stage<T, PHI, DT, J, NB, S>), dim3(nb), dim3(NB), lds, s, m, tm, tv, hv, pp, (int)grad,
                           (int)lim, bcs, qin, q0, qout, bincr, dt, dt_inst, alpha, res_blk, ctrl, rec);
        return;
    }
    if (resident == 0 || resident_lds != lds) {
        // workgroups per CU by the 160 KB of LDS (with 1 KB for the static
        // arrays) and by the 32 wave slots.  The persistent form's registers
        // (about 150 VGPRs: three waves per SIMD) can allow fewer; more
        // workgroups than fit at once only queue behind the resident ones
        // (none waits on another).
        int dev = 0, cus = 0;
        const int per_cu = std::max(1, std::min((int)(160 * 1024 / (lds + 1024)), 32 / (NB / 64)));
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = kXcd;
        (void)hipGetLastError();
        resident = std::max(kXcd, per_cu * cus / kXcd * kXcd);
        resident_lds = lds;
    }
    const int nb = std::min(resident, xcd_grid_blocks(tm.n_rt));
    hipLaunchKernelGGL((k_stage_q<T, PHI, DT, J, NB, S>), dim3(nb), dim3(NB), lds, s, m, tm, tv, hv, pp, (int)grad,
                       (int)lim, bcs, qin, q0, qout, bincr, dt, dt_inst, alpha, res_blk, ctrl, rec, queue);
}

// Workgroup size: a tile per workgroup either way.  With 256 lanes the
// phases loop over a tile's cells and edges, and four workgroups share a
// CU when their LDS fits (state registers allow 16 waves per CU): every
// wave has work in every phase.  Three workgroups per CU still beat two
// of 512 lanes when a tile and its ring fit one round of the gradient
// phase; larger halos run with 512 lanes, two workgroups per CU
// (stage_lanes).
template <class T, int PHI, int DT, bool J, class S>
bool launch_stage_k(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const TileView& tv, const HaloView& hv,
                    const PhysParams& pp, bool grad, bool lim, const BcParam<T>* bcs, const Q4<S>* qin,
                    const Q4<S>* q0, Q4<S>* qout, Q4<T>* bincr, T* dt, const T* dt_inst, T alpha,
                    double* res_blk, const Ctrl* ctrl, int threads, const StageRecord<T>& rec, int32_t* queue) {
    const size_t lds = stage_lds_bytes<T, S>(hv);
    if (stage_lanes(lds, threads, hv.mean_g) == kStageBlock / 2)
        launch_stage_nb<T, PHI, DT, J, kStageBlock / 2, S>(s, m, tm, tv, hv, pp, grad, lim, bcs, qin, q0, qout,
                                                           bincr, dt, dt_inst, alpha, res_blk, ctrl, lds, rec, queue);
    else
        launch_stage_nb<T, PHI, DT, J, kStageBlock, S>(s, m, tm, tv, hv, pp, grad, lim, bcs, qin, q0, qout, bincr,
                                                       dt, dt_inst, alpha, res_blk, ctrl, lds, rec, queue);
    return true;
}

template <class T, class S>
bool launch_stage(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const TileView& tv, const HaloView& hv,
                  const PhysParams& pp, int phi_kind, bool joined, bool grad, bool lim, int dt_kind,
                  const BcParam<T>* bcs, const Q4<S>* qin, const Q4<S>* q0, Q4<S>* qout, Q4<T>* bincr, T* dt,
                  const T* dt_inst, T alpha, double* res_blk, const Ctrl* ctrl, int threads,
                  const StageRecord<T>& rec, int32_t* queue) {
    if (tm.n_rt == 0) return true;
    if (stage_lds_bytes<T, S>(hv) > 160 * 1024) return false;
#define FVH_ST(P, D, J) launch_stage_k<T, P, D, J, S>(s, m, tm, tv, hv, pp, grad, lim, bcs, qin, q0, qout, bincr, \
                                                     dt, dt_inst, alpha, res_blk, ctrl, threads, rec, queue)
#define FVH_STJ(P, D) (joined ? FVH_ST(P, D, true) : FVH_ST(P, D, false))
#define FVH_STD(P) (dt_kind == 1 ? FVH_STJ(P, 1) : dt_kind == 2 ? FVH_STJ(P, 2) : FVH_STJ(P, 0))
    switch (phi_kind) {
        case PHI_ZERO: return FVH_STD(PHI_ZERO);
        case PHI_MICHALAK: return FVH_STD(PHI_MICHALAK);
        case PHI_VENKAT: return FVH_STD(PHI_VENKAT);
        case PHI_VENKAT_S:
            // smoothed kinks: FP64 only, the time step from the time-step kernel
            if constexpr (kSmoothable<T, S>) return FVH_STJ(PHI_VENKAT_S, 0);
            return false;
        default: return FVH_STD(PHI_MINMOD);
    }
#undef FVH_STD
#undef FVH_STJ
#undef FVH_ST
}

template <class T, class S>
void launch_bounds_init(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, bool lin,
                        const BcParam<T>* bcs, Q4<S>* q, const Q4<T>* gx, const Q4<T>* gy, const Q4<T>* lim) {
    const int nb = nblocks(m.n_real);
    if (nb == 0) return;
    hipLaunchKernelGGL((k_bounds_init<T, S>), dim3(nb), dim3(kBlock), 0, s, m, pp, (int)lin, bcs, q, gx, gy, lim);
}

template <class T>
void launch_phase_limiters(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, int phi_kind, const Q4<T>* q,
                           const Q4<T>* gx, const Q4<T>* gy, Q4<T>* lim, Q4<T>* qmin, Q4<T>* qmax) {
    const int nb = nblocks(m.n_total);
    if (nb == 0) return;
#define FVH_LM(P) hipLaunchKernelGGL((k_limiters<T, P>), dim3(nb), dim3(kBlock), 0, s, m, pp, q, gx, gy, lim, qmin, qmax)
    switch (phi_kind) {
        case PHI_ZERO: FVH_LM(PHI_ZERO); break;
        case PHI_MICHALAK: FVH_LM(PHI_MICHALAK); break;
        case PHI_VENKAT: FVH_LM(PHI_VENKAT); break;
        default: FVH_LM(PHI_MINMOD); break;
    }
#undef FVH_LM
}

template <class T>
void launch_phase_faces(hipStream_t s, const MeshView<T>& m, const Q4<T>* q, const Q4<T>* gx, const Q4<T>* gy,
                        const Q4<T>* lim, Q4<T>* face) {
    const int nb = nblocks(m.n_total);
    if (nb == 0) return;
    hipLaunchKernelGGL((k_faces<T>), dim3(nb), dim3(kBlock), 0, s, m, q, gx, gy, lim, face);
}

void launch_update_cells(hipStream_t s, double* q, const double* ql, const double* qt, const double* dt, double v,
                         int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_update_cells, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, q, ql, qt,
                       dt, v, n);
}

int array_min_blocks(int64_t n) {
    const int64_t nb = (n + kBlock - 1) / kBlock;
    return (int)(nb < kArrayMinBlocks ? (nb < 1 ? 1 : nb) : kArrayMinBlocks);
}

void launch_array_min(hipStream_t s, double* a, int64_t n, double* part) {
    if (n <= 0) return;
    const int nb = array_min_blocks(n);
    hipLaunchKernelGGL(k_array_min, dim3(nb), dim3(kBlock), 0, s, (const double*)a, n, part);
    hipLaunchKernelGGL(k_array_min_fill, dim3(1), dim3(kBlock), 0, s, a, n, (const double*)part, nb);
}

template <class T>
void launch_residuals(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const InstView& iv, const Q4<T>* qt,
                      double* blk, double* out) {
    if (tm.n_rt == 0 || iv.n == 0) return;
    hipLaunchKernelGGL((k_res_tile<T>), dim3(tm.n_rt), dim3(kBlock), 0, s, m, tm, qt, blk);
    hipLaunchKernelGGL(k_res_inst, dim3(iv.n), dim3(kBlock), 0, s, iv, (const double*)blk, out);
}

void launch_ctrl_begin(hipStream_t s, int n_inst, Ctrl* ctrl, uint32_t max_step, double max_time, double tol,
                       int stop_var) {
    if (n_inst == 0) return;
    hipLaunchKernelGGL(k_ctrl_begin, dim3(nblocks(n_inst)), dim3(kBlock), 0, s, n_inst, ctrl, max_step, max_time,
                       tol, stop_var);
}

void launch_ctrl_end(hipStream_t s, const InstView& iv, Ctrl* ctrl, const double* res_blk, int need_res,
                     const void* dt_src, int dt_is_f32, int dt_is_inst) {
    if (iv.n == 0) return;
    if (dt_is_f32)
        hipLaunchKernelGGL((k_ctrl_end<float>), dim3(iv.n), dim3(kBlock), 0, s, iv, ctrl, res_blk, need_res,
                           (const float*)dt_src, dt_is_inst);
    else
        hipLaunchKernelGGL((k_ctrl_end<double>), dim3(iv.n), dim3(kBlock), 0, s, iv, ctrl, res_blk, need_res,
                           (const double*)dt_src, dt_is_inst);
}

template <class T, class S>
void launch_sums(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const InstView& iv, const Q4<S>* q,
                 double* blk, double* out) {
    if (tm.n_rt == 0 || iv.n == 0) return;
    hipLaunchKernelGGL((k_sums_tile<T, S>), dim3(tm.n_rt), dim3(kBlock), 0, s, m, tm, q, blk);
    hipLaunchKernelGGL(k_sums_inst, dim3(iv.n), dim3(kBlock), 0, s, iv, (const double*)blk, out);
}

#define FVH_INST(T, S)                                                                                       \
    template void launch_dt<T, S>(hipStream_t, const MeshView<T>&, const TileMap&, int, const PhysParams&,     \
                                  const Q4<S>*, T*, double*, const Ctrl*);                                   \
    template void launch_gradlim<T, S>(hipStream_t, const MeshView<T>&, const TileMap&, const PhysParams&,     \
                                       int, bool, bool, bool, const Q4<S>*, Q4<T>*, Q4<T>*, int, T*, double*, \
                                       Q4<T>*, Q4<T>*, Q4<T>*, const Ctrl*);                                 \
    template void launch_flux<T, S>(hipStream_t, const MeshView<T>&, const TileMap&, const TileView*,         \
                                    const PhysParams&, int, bool, const BcParam<T>*, const Q4<S>*,            \
                                    const Q4<S>*, Q4<S>*, const Q4<T>*, const Q4<T>*, const T*, const T*, T,  \
                                    double*, Q4<T>*, const Ctrl*, Q4<T>*, const Q4<T>*, const Q4<T>*);        \
    template void launch_smooth<T, S>(hipStream_t, const MeshView<T>&, const TileMap&, const PhysParams&,     \
                                      bool, const BcParam<T>*, const Q4<T>*, const Q4<T>*, Q4<T>*,            \
                                      const Q4<S>*, Q4<S>*, const Q4<T>*, const T*, const T*, T, double*,     \
                                      Q4<T>*, const Ctrl*);                                                  \
    template void launch_bounds_init<T, S>(hipStream_t, const MeshView<T>&, const PhysParams&, bool,         \
                                           const BcParam<T>*, Q4<S>*, const Q4<T>*, const Q4<T>*,            \
                                           const Q4<T>*);                                                    \
    template void launch_sums<T, S>(hipStream_t, const MeshView<T>&, const TileMap&, const InstView&,        \
                                    const Q4<S>*, double*, double*);                                         \
    template bool launch_stage<T, S>(hipStream_t, const MeshView<T>&, const TileMap&, const TileView&,       \
                                     const HaloView&, const PhysParams&, int, bool, bool, bool, int,         \
                                     const BcParam<T>*, const Q4<S>*, const Q4<S>*, Q4<S>*, Q4<T>*, T*,     \
                                     const T*, T, double*, const Ctrl*, int, const StageRecord<T>&,          \
                                     int32_t*);                                                              \
    template size_t stage_lds_bytes<T, S>(const HaloView&);
FVH_INST(double, double)
FVH_INST(float, float)
FVH_INST(double, float)
template void launch_phase_limiters<double>(hipStream_t, const MeshView<double>&, const PhysParams&, int,
                                            const Q4<double>*, const Q4<double>*, const Q4<double>*, Q4<double>*,
                                            Q4<double>*, Q4<double>*);
template void launch_phase_faces<double>(hipStream_t, const MeshView<double>&, const Q4<double>*, const Q4<double>*,
                                         const Q4<double>*, const Q4<double>*, Q4<double>*);
template void launch_residuals<double>(hipStream_t, const MeshView<double>&, const TileMap&, const InstView&,
                                       const Q4<double>*, double*, double*);

}  // namespace fvh
