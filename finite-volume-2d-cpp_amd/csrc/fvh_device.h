// fvh_device.h -- device helpers shared by the forward and the adjoint
// kernel files (block reductions, XCD-aware block order, state checks).
#pragma once

#include "fvh_kernels.h"

namespace fvh {
namespace {

constexpr double kInf = __builtin_inf();

__device__ inline bool stopped(const Ctrl* ctrl, int inst) { return ctrl != nullptr && ctrl[inst].stop != 0; }

// NaN-ignoring minimum over a 256-thread block (std::min(a,b) keeps a unless
// b < a, so NaN never wins once a finite value is held).
__device__ inline double block_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = smin(v, __shfl_xor(v, off, 64));
    __shared__ double part[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    double r = part[0];
    for (int w = 1; w < kBlock / 64; ++w) r = smin(r, part[w]);
    return r;
}

// Fixed-pattern sum of 4 values over an NB-thread block (deterministic).
template <int NB = kBlock>
__device__ inline void block_sum4(double* v) {
    for (int k = 0; k < 4; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    __shared__ double part[NB / 64][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 4; ++k) part[wave][k] = v[k];
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
        double s = part[0][k];
        for (int w = 1; w < NB / 64; ++w) s += part[w][k];
        v[k] = s;
    }
}

template <class T>
__device__ inline Q4<T> ld(const Q4<T>* p) { return *p; }

// State loads and stores between the storage type S of the cell states and
// the arithmetic type T (identical except in the FP32-state, FP64-arithmetic
// mode, where a load widens exactly and a store rounds to nearest).
template <class T, class S>
__device__ inline Q4<T> ldq(const Q4<S>* p) {
    const Q4<S> a = *p;
    Q4<T> r;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.v[k] = T(a.v[k]);
    return r;
}
template <class T, class S>
__device__ inline void stq(Q4<S>* p, const Q4<T>& a) {
    Q4<S> r;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.v[k] = S(a.v[k]);
    *p = r;
}

// XCD-aware block order.  The dispatcher hands workgroup b to XCD b % 8, and
// each XCD has its own L2; remapping hardware block b to logical block
// (b % 8) * per + b / 8 gives every XCD one contiguous range of cells, so a
// cell's neighbours (close in the renumbered order) sit in the same L2.
constexpr int kXcd = 8;
// half-edges of one cell whose loads the per-cell loops issue together
constexpr int kHeChunk = 3;
__device__ inline int xcd_block(int nblk) {
    const int per = (nblk + kXcd - 1) / kXcd;
    const int b = blockIdx.x;
    const int lb = (b % kXcd) * per + b / kXcd;
    return lb < nblk ? lb : -1;
}

template <class T>
__device__ inline bool physical(T g, const T* q) {
    return q[0] > T(0) && pressure(g, q) > T(0);
}

// Limited linear increment (grad q . d) phi at an edge centre (explicit.cpp:307-313).
template <class T>
__device__ inline Q4<T> face_inc(const T (&gx)[4], const T (&gy)[4], const T (&lim)[4], const G2<T>& ec,
                                 const G2<T>& cc) {
    const T dx = ec.v[0] - cc.v[0];
    const T dy = ec.v[1] - cc.v[1];
    Q4<T> inc;
#pragma unroll
    for (int k = 0; k < 4; ++k) inc.v[k] = (gx[k] * dx + gy[k] * dy) * lim[k];
    return inc;
}

inline int nblocks(int n) { return (n + kBlock - 1) / kBlock; }
// hardware grid for xcd_block(): a multiple of the XCD count
inline int xcd_grid_blocks(int nb) { return ((nb + kXcd - 1) / kXcd) * kXcd; }
inline int xcd_grid(int n) { return xcd_grid_blocks(nblocks(n)); }

}  // namespace
}  // namespace fvh
