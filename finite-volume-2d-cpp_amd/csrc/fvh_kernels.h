// fvh_kernels.h -- gfx950 kernels of the explicit residual-and-update path.
//
// Data layout in HBM (device cell order: a locality renumbering of the real
// cells -- recursive-bisection tiles or reverse Cuthill-McKee -- then the
// boundary ghost cells in boundary-element order):
//   state      Q4<T>[cells]          AoS, 4 conserved variables per cell, so a
//                                    neighbour gather is one 32 B (FP64) read
//   half-edges CSR per cell, sorted by reference edge id and side, so every
//              per-cell sum runs in the reference's summation order:
//              he_off[cells+1], he_nbr, he_edge (edge | bc<<30 | side<<31),
//              he_twin, he_bc (BC slot of a boundary edge's interior side),
//              he_slot (the edge's index in its tile's edge list)
//   edges      G4<T>(nx, ny, length, Green-Gauss weight) + G2<T>(centre)
//   cells      area, 1/area, sqrt(area), centre
//   faces      Q4<T>[half-edges]  limited linear reconstruction
//              q + (grad q . d) phi at each half-edge, written by the
//              gradient/limiter pass, read by the flux pass
//   bincr      Q4<T>[ghosts]      the increment (grad q . d) phi of each
//              boundary edge's interior side, for the boundary update
//   tiles      per 256-cell tile: the edges it touches, each with its two
//              half-edges and cells
//
// No atomics and no colouring: every per-cell sum is a gather in the
// reference's edge order, so results are deterministic and bit-identical to
// the reference CPU run in FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "fvh_physics.h"

namespace fvh {

template <class T> struct alignas(4 * sizeof(T)) Q4 { T v[4]; };
template <class T> struct alignas(2 * sizeof(T)) G2 { T v[2]; };

constexpr uint32_t kSideBit = 0x80000000u;
constexpr uint32_t kBcBit = 0x40000000u;
constexpr uint32_t kEdgeMask = 0x3fffffffu;
constexpr int kMaxHe = 4;   // half-edges of a real cell the adjoint takes (triangles and quadrilaterals)
constexpr int kBlock = 256;        // cells per workgroup = cells per tile
constexpr int kTileEdges = 1024;   // LDS edge slots per tile
constexpr int kStageBlock = 512;   // largest workgroup of the fused stage pass (one tile per workgroup)
constexpr int kMaxBc = 16;

// Per-run control block, advanced on the device so that steps can be queued
// without host round trips; every kernel is a no-op once `stop` is set.
struct Ctrl {
    double time;
    double time_prev;  // time before the last step
    double dt0;        // dt of reference cell 0 in the current step
    double R[4];
    double R0[4];
    uint32_t step;
    uint32_t stop;
    uint32_t nonfinite;
    uint32_t pad;
};

template <class T>
struct MeshView {
    int32_t n_real, n_total;
    const int32_t* he_off;
    const int32_t* he_nbr;
    const uint32_t* he_edge;
    const int32_t* he_twin;
    const int32_t* he_bc;
    const uint16_t* he_slot;
    const Q4<T>* e_geo;   // nx, ny, len, gg weight
    const G2<T>* e_c;     // edge centre
    const T* c_area;
    const T* c_inv_area;
    const T* c_sqrt_area;
    const G2<T>* c_c;     // cell centre
};

// Tiles of at most kBlock consecutive device cells, each owned by one
// instance (a single mesh is a batch of one instance).  The tiles of real
// cells come first, [0, n_rt), instance by instance; then the tiles of the
// boundary ghost cells, [n_rt, n_all).  One workgroup processes one tile.
struct TileMap {
    int32_t n_rt, n_all;
    const int32_t* c0;     // [n_all+1] first device cell of each tile
    const int32_t* inst;   // [n_all] instance owning each tile
};

// Per-instance ranges for the segmented reductions (time-step minimum,
// residual sums) and the run bookkeeping.
struct InstView {
    int32_t n;
    const int4* tiles;     // (rt0, rt1, gt0, gt1): real and ghost tile ranges
    const int32_t* dev0;   // device position of the instance's reference cell 0
};

// Edge lists of the real-cell tiles for the tiled flux pass.
struct TileView {
    int32_t n_tiles;
    const int32_t* off;    // [n_tiles+1] into the tile edge lists
    const int32_t* edge;   // device edge id
    const int2* he;        // (half-edge of side 0, half-edge of side 1)
    const int2* cell;      // (cell of side 0, cell of side 1)
};

// Halo of each real tile for the fused stage pass: the tile's ring-1
// neighbours (real cells first, then ghosts) and the ring-2 neighbours of its
// real ring-1 cells, with every neighbour reference pre-resolved to an index
// into the tile's local cell list [tile cells | ring 1 | ring 2].
struct HaloView {
    const int4* ext;          // [n_rt] (offset into cells, ring-1 real, ring-1 total, ring-2)
    const int32_t* cells;     // ring-1 then ring-2 device cells of each tile
    const int32_t* r1_start;  // parallel to cells: a real ring-1 cell's first entry in r1_loc
    const uint16_t* r1_loc;   // local neighbour index per half-edge of real ring-1 cells
    const uint16_t* he_loc;   // per half-edge of a real cell: local neighbour index in its own tile
    const uint16_t* r1_slot;  // parallel to r1_loc: the tile's edge slot of that half-edge (0xffff: none)
    int32_t max_ext, max_g, max_edges;   // largest tile lists
    int32_t mean_g;                      // tile plus real ring-1 cells, averaged over tiles
    int32_t lds_bytes;                   // largest tile's LDS (stage_lds_tile); 0: from the maxima above
};

struct PhysParams {
    double gamma, cfl, dt_fixed, K, phi_yt;
    double mu, pr, cp, r;   // laminar viscous constants (viscous flux and time steps)
    int positivity;   // first-order fallback for non-physical face states
    int diffusive;    // face gradients for the viscous flux
    // widths of the smoothed kinks (FVH_FLAG_SMOOTH_KINKS): per conserved
    // variable (limiter extrema and the sign blend of the gradient
    // projection), wave speeds (entropy cure, summed-CFL time step), and the
    // limiter's running minimum (dimensionless)
    double kw[4], kwv, kwl;
    // per instance: set to 1 by any stage update that leaves a real cell with
    // rho <= 0, p <= 0 or a non-finite value (null: no check)
    uint32_t* unphysical;
};

// Launchers (implemented in fvh_kernels.hip).  All take the solver stream;
// ctrl is the per-instance control array (nullptr: no stop test).  T is the
// arithmetic type (and that of the geometry and of every intermediate
// array), S the storage type of the cell states: (double, double) and
// (float, float), and (double, float) for FP32 states with FP64 arithmetic.
template <class T, class S>
void launch_dt(hipStream_t s, const MeshView<T>& m, const TileMap& tm, int dt_kind, const PhysParams& pp,
               const Q4<S>* q, T* dt, double* blk_min, const Ctrl* ctrl);
// Per-instance minimum of the time step over real and ghost cells (min_dt).
void launch_dt_min(hipStream_t s, const InstView& iv, const double* blk_min, const void* dt_cells,
                   int dt_is_f32, void* dt_global, const Ctrl* ctrl);
// Smoothed kinks: replaces each instance's minimum in dt_global (from
// launch_dt_min) by the p-norm (sum_i dt_i^-128)^(-1/128) of its cells'
// steps (FP64; part: two words per tile).
void launch_dt_soft(hipStream_t s, const TileMap& tm, const InstView& iv, const double* dt_cells, double* dt_global,
                    uint64_t* part, const Ctrl* ctrl);
// The kernels with smoothed kinks (internal kinds PHI_VENKAT_S, flux 3 and 4,
// time step 5) exist for FP64 states and arithmetic only.
template <class T, class S>
constexpr bool kSmoothable = std::is_same<T, double>::value && std::is_same<S, double>::value;
// Gradient + limiter + face states; with dt_kind >= 0 also the time step of
// every cell and per-tile minima (first stage of a step).  dbg_gx/dbg_gy
// (when non-null) receive every cell's gradient, zero for ghosts; dbg_lim
// (when non-null) its limiter.
template <class T, class S>
void launch_gradlim(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const PhysParams& pp, int phi_kind,
                    bool grad, bool lim, bool lin, const Q4<S>* q, Q4<T>* face, Q4<T>* bincr,
                    int dt_kind, T* dt, double* blk_min,
                    Q4<T>* dbg_gx, Q4<T>* dbg_gy, Q4<T>* dbg_lim, const Ctrl* ctrl);
// Flux + RK update + boundary update; tiles == nullptr selects the per-cell
// form that computes every interior edge flux from both sides.  dt_inst
// (global time step per instance) replaces the per-cell dt when non-null.
// flux_kind: FVH_FLUX_* (0 split, 1 joined, 2 split + viscous terms; the
// viscous kind reads the cell gradients gxc, gyc when pp.diffusive).
template <class T, class S>
void launch_flux(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const TileView* tiles,
                 const PhysParams& pp, int flux_kind, bool lin, const BcParam<T>* bcs, const Q4<S>* qin,
                 const Q4<S>* q0, Q4<S>* qout, const Q4<T>* face, const Q4<T>* bincr, const T* dt,
                 const T* dt_inst, T alpha, double* res_blk, Q4<T>* dbg_qt, const Ctrl* ctrl,
                 Q4<T>* qt_only = nullptr, const Q4<T>* gxc = nullptr, const Q4<T>* gyc = nullptr);
// One sweep of implicit residual smoothing (explicit.cpp:32-66) over the qt
// that a flux pass with qt_only stored: qt_out != nullptr stores the sweep's
// qt; qt_out == nullptr finishes the stage from it (update, boundary update,
// residual sums), taking the flux pass's place.
template <class T, class S>
void launch_smooth(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const PhysParams& pp, bool lin,
                   const BcParam<T>* bcs, const Q4<T>* qt_nb, const Q4<T>* qt0, Q4<T>* qt_out, const Q4<S>* q0,
                   Q4<S>* qout, const Q4<T>* bincr, const T* dt, const T* dt_inst, T alpha, double* res_blk,
                   Q4<T>* dbg_qt, const Ctrl* ctrl);
// Optional outputs of a fused stage for the adjoint's replays: every cell's
// gradient and limiter (ghosts: zero gradient, reset limiter) and every real
// cell's qt, as the two-pass form stores them.  All null: none.
template <class T>
struct StageRecord {
    Q4<T>* gx = nullptr;
    Q4<T>* gy = nullptr;
    Q4<T>* lm = nullptr;
    Q4<T>* qt = nullptr;
};

// Fused stage: gradient + limiter (tile and ring 1, from LDS), face states,
// edge fluxes, gather, RK update and boundary update in one pass per tile.
// dt_kind >= 0 also computes the tile cells' local time step (first stage).
// threads: workgroup size, 256 or 512; 0 picks 256 when four workgroups'
// halos fit a CU's LDS, else 512.  queue: null for one workgroup per tile;
// else kXcd + 1 zeroed counters (left zeroed) of the persistent form, whose
// resident workgroups pull the tiles from per-XCD queues.  Both forms give
// the same bits.
// Returns false when the tile halo does not fit the workgroup's LDS.
template <class T, class S>
bool launch_stage(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const TileView& tv, const HaloView& hv,
                  const PhysParams& pp, int phi_kind, bool joined, bool grad, bool lim, int dt_kind,
                  const BcParam<T>* bcs, const Q4<S>* qin, const Q4<S>* q0, Q4<S>* qout, Q4<T>* bincr, T* dt,
                  const T* dt_inst, T alpha, double* res_blk, const Ctrl* ctrl, int threads = 0,
                  const StageRecord<T>& rec = StageRecord<T>(), int32_t* queue = nullptr);
// LDS of the fused stage pass for a tile of n_ext local cells and n_edges
// edges: the states of its local cell list in the storage type, then, from
// the next 32-byte boundary, two face states per tile edge in the
// arithmetic type, then (positivity fallback on) the pressures of those
// face states.  A launch takes the largest tile's (stage_lds_bytes).
template <class S>
__host__ __device__ inline size_t stage_lds_states(int32_t n_ext) {
    return ((size_t)n_ext * sizeof(Q4<S>) + 31) / 32 * 32;
}
template <class T, class S>
__host__ __device__ inline size_t stage_lds_tile(int32_t n_ext, int32_t n_edges, bool pressures = false) {
    return stage_lds_states<S>(n_ext) + (size_t)(2 * n_edges) * sizeof(Q4<T>) +
           (pressures ? (size_t)(2 * n_edges) * sizeof(T) : 0);
}
template <class T, class S>
size_t stage_lds_bytes(const HaloView& hv);
// Workgroup size the fused stage pass takes for a halo of lds bytes when
// asked for `threads`.  0: 256 lanes when four workgroups fit a CU's LDS, or
// when three fit and an average tile with its ring-1 cells (mean_g) fits the
// lanes, so that the gradient phase takes one round; 512 otherwise.
inline int stage_lanes(size_t lds, int threads, int mean_g) {
    if (threads) return threads;
    if (lds + 1024 <= 160 * 1024 / 4) return kStageBlock / 2;
    if (lds + 1024 <= 160 * 1024 / 3 && mean_g <= kStageBlock / 2) return kStageBlock / 2;
    return kStageBlock;
}
// Boundary update of the ghost cells from the real cells' states, with
// given gradients and limiters (null: zero ones, the initial update).
template <class T, class S>
void launch_bounds_init(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, bool lin,
                        const BcParam<T>* bcs, Q4<S>* q, const Q4<T>* gx = nullptr, const Q4<T>* gy = nullptr,
                        const Q4<T>* lim = nullptr);
// The phases of explicit.cpp one at a time on given arrays (device order):
// limiters, neighbour minima and maxima from given gradients; face states
// from given gradients and limiters (for launch_flux with qt_only);
// update_cells over n values; min_dt over n values in place (part: at least
// array_min_blocks(n) doubles); residual norms per instance from a given qt.
template <class T>
void launch_phase_limiters(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, int phi_kind, const Q4<T>* q,
                           const Q4<T>* gx, const Q4<T>* gy, Q4<T>* lim, Q4<T>* qmin, Q4<T>* qmax);
template <class T>
void launch_phase_faces(hipStream_t s, const MeshView<T>& m, const Q4<T>* q, const Q4<T>* gx, const Q4<T>* gy,
                        const Q4<T>* lim, Q4<T>* face);
void launch_update_cells(hipStream_t s, double* q, const double* ql, const double* qt, const double* dt, double v,
                         int64_t n);
constexpr int kArrayMinBlocks = 1024;
constexpr int64_t kMinDtDeviceMax = (int64_t)1 << 31;   // values fvh_min_dt_device accepts
int array_min_blocks(int64_t n);
void launch_array_min(hipStream_t s, double* a, int64_t n, double* part);
template <class T>
void launch_residuals(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const InstView& iv, const Q4<T>* qt,
                      double* blk, double* out);
void launch_ctrl_begin(hipStream_t s, int n_inst, Ctrl* ctrl, uint32_t max_step, double max_time, double tol,
                       int stop_var);
void launch_ctrl_end(hipStream_t s, const InstView& iv, Ctrl* ctrl, const double* res_blk, int need_res,
                     const void* dt_src, int dt_is_f32, int dt_is_inst);
// Adjoint of the mesh geometry, summed over every reversed stage and step:
// per half-edge (n_x, n_y, length, Green-Gauss weight) and (edge centre x, y),
// written only by the thread of the half-edge's cell; per edge (n_x, n_y,
// length, unused) from the edge pass; per cell (area, centre x, y, unused).
// With he == nullptr no geometry adjoint is formed.
template <class T>
struct GeoAdj {
    Q4<T>* he = nullptr;
    G2<T>* hc = nullptr;
    Q4<T>* e = nullptr;
    Q4<T>* c = nullptr;
};

// Discrete adjoint (fvh_adjoint.hip): the reverse of one RK stage, of the
// global time step's minimum cell, of the initial boundary update, and the
// boundary-pressure objective with its seed.  ga: geometry adjoint (or none).
// lam_out == nullptr: the adjoint of the stage output is gathered from the
// parts (lam_direct, lam_face, nb_contrib) the previous call left; lam_in ==
// nullptr: this call's parts are left for the next call instead of gathered.
template <class T>
void launch_adj_stage(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const PhysParams& pp, int phi_kind,
                      bool grad, bool lim, bool lin, const BcParam<T>* bcs, int n_edges, const int4* epair,
                      const Q4<T>* qin, const Q4<T>* qout, const Q4<T>* gx, const Q4<T>* gy, const Q4<T>* lm,
                      const Q4<T>* qt, const T* dt, const T* dt_g, T alpha, const Q4<T>* lam_out, Q4<T>* lam_q0,
                      Q4<T>* lam_qt, Q4<T>* lam_bq, Q4<T>* lam_bincr, Q4<T>* lam_face, Q4<T>* lam_direct,
                      Q4<T>* nb_contrib, Q4<T>* lam_in, int accumulate_in, T* lam_dt, double* ld_blk,
                      const GeoAdj<T>& ga);
constexpr int kArgminBlocks = 256;
template <class T>
void launch_dt_argmin(hipStream_t s, const T* dt, const int32_t* ref, int n, double* bv, int32_t* bref,
                      int32_t* bdev, int32_t* out);
template <class T>
void launch_adj_dt(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, int dt_kind, const Q4<T>* q,
                   const int32_t* cell, const double* ld_blk, int n_blk, Q4<T>* lam_q0, const GeoAdj<T>& ga);
// Smoothed kinks: the reverse of the p-norm global time step over every
// cell (ld_blk: the per-tile partials of its adjoint, summed into lam_dt;
// nb_dt: one state per half-edge of scratch).
template <class T>
void launch_adj_dt_soft(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, const Q4<T>* q, const T* dt,
                        const T* dt_g, const double* ld_blk, int n_blk, double* lam_dt, Q4<T>* lam_q0, Q4<T>* nb_dt,
                        const GeoAdj<T>& ga);
template <class T>
void launch_adj_bounds_init(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, const BcParam<T>* bcs,
                            const Q4<T>* q, Q4<T>* lam, const GeoAdj<T>& ga);
// terms[b] = p_b (w.n) len per boundary element; p_out (optional) = p_b
template <class T>
void launch_objective(hipStream_t s, int n_b, const PhysParams& pp, const int32_t* ghost, const Q4<T>* geo_w,
                      const Q4<T>* q, double* terms, Q4<T>* lam, double* p_out);
template <class T>
void launch_fill(hipStream_t s, Q4<T>* a, int n);
// n states from a to b (a kernel: cheaper to issue than the runtime's copy)
template <class T>
void launch_copy(hipStream_t s, const void* a, void* b, int n);

// FP64 conserved sums sum_i A_i q_ik over each instance's real cells.
template <class T, class S>
void launch_sums(hipStream_t s, const MeshView<T>& m, const TileMap& tm, const InstView& iv, const Q4<S>* q,
                 double* blk, double* out);

}  // namespace fvh
