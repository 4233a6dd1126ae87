// fvh_kernels.h -- gfx950 kernels of the explicit residual-and-update path.
//
// Data layout in HBM (device cell order: reverse Cuthill-McKee of the real
// cells, then the boundary ghost cells in boundary-element order):
//   state      Q4<T>[cells]          AoS, 4 conserved variables per cell, so a
//                                    neighbour gather is one 32 B (FP64) read
//   half-edges CSR per cell, sorted by reference edge id and side, so every
//              per-cell sum runs in the reference's summation order:
//              he_off[cells+1], he_nbr, he_edge (edge | side<<31), he_twin,
//              he_bc (BC slot on the interior side of a boundary edge, else -1)
//   edges      G4<T>(nx, ny, length, Green-Gauss weight) + G2<T>(centre)
//   cells      area, 1/area, sqrt(area), centre
//   increments Q4<T>[half-edges]  limited reconstruction (grad.d)*phi per
//              half-edge, written by the gradient+limiter pass and read by the
//              flux pass (own and twin side) and the boundary update
//
// Every kernel is one thread per cell and gathers over the cell's half-edges:
// no atomics, no colouring, deterministic, and the order of every
// floating-point sum equals the reference's edge loop order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fvh_physics.h"

namespace fvh {

template <class T> struct alignas(4 * sizeof(T)) Q4 { T v[4]; };
template <class T> struct alignas(2 * sizeof(T)) G2 { T v[2]; };

constexpr uint32_t kSideBit = 0x80000000u;
constexpr uint32_t kEdgeMask = 0x7fffffffu;
constexpr int kBlock = 256;
constexpr int kMaxBc = 16;

// Per-run control block, advanced on the device so that steps can be queued
// without host round trips; every kernel is a no-op once `stop` is set.
struct Ctrl {
    double time;
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
    const Q4<T>* e_geo;   // nx, ny, len, gg weight
    const G2<T>* e_c;     // edge centre
    const T* c_area;
    const T* c_inv_area;
    const T* c_sqrt_area;
    const G2<T>* c_c;     // cell centre
};

template <class T>
struct BcTable {
    BcParam<T> p[kMaxBc];
};

struct PhysParams {
    double gamma, cfl, dt_fixed, K, phi_yt;
    int positivity;   // first-order fallback for non-physical face states
};

// Launchers (implemented in fvh_kernels.hip).  All take the solver stream.
template <class T>
void launch_dt(hipStream_t s, const MeshView<T>& m, int dt_kind, const PhysParams& pp,
               const Q4<T>* q, T* dt, double* blk_min, const Ctrl* ctrl);
void launch_dt_min(hipStream_t s, const double* blk_min, int nblk, const void* dt_cells,
                   int dt_is_f32, int32_t dev0, double* dt_global, const Ctrl* ctrl);
template <class T>
void launch_gradlim(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, int phi_kind,
                    bool grad, bool lim, bool lin, const Q4<T>* q, Q4<T>* incr,
                    Q4<T>* dbg_gx, Q4<T>* dbg_gy, Q4<T>* dbg_lim, const Ctrl* ctrl);
template <class T>
void launch_flux(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, bool joined, bool lin,
                 const BcTable<T>& bcs, const Q4<T>* qin, const Q4<T>* q0, Q4<T>* qout,
                 const Q4<T>* incr, const T* dt, int dt_stride, const double* dt_global, T alpha,
                 double* res_blk, Q4<T>* dbg_qt, const Ctrl* ctrl);
template <class T>
void launch_bounds_init(hipStream_t s, const MeshView<T>& m, const PhysParams& pp, bool lin,
                        const BcTable<T>& bcs, Q4<T>* q);
void launch_ctrl_begin(hipStream_t s, Ctrl* ctrl, uint32_t max_step, double max_time, double tol);
void launch_ctrl_end(hipStream_t s, Ctrl* ctrl, const double* res_blk, int nblk, int need_res,
                     const void* dt_src, int dt_is_f32, int32_t dt_index);

}  // namespace fvh
