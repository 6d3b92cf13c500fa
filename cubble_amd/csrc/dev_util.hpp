// dev_util.hpp -- device helpers shared by the kernels: block context,
// XCD-aware block remap, deterministic wave64 / block reductions, block
// scans, and the geometry (minimum-image difference, cell grid) whose
// rounding must match the FP64 restatement bit for bit.  Compiled with
// -ffp-contract=off: no fused multiply-add.
//
// No kernel communicates between its own blocks: the eight XCD L2 caches of
// MI355X are not coherent with each other, so a cross-block hand-off inside
// one launch needs an agent-scope release (an L2 write-back) per block.
// Reductions across blocks happen in a following one-block-per-foam kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "foam.hpp"

namespace cubble_amd {

struct Ctx {
    Foam *F;
    int lb;  // logical block within the foam
};

// Workgroup g of a launch runs on XCD g % 8.  Blocks of one residue class
// get one contiguous range of the foam, so each XCD's L2 holds a contiguous
// slab of the cell-ordered bubbles and the neighbours they gather.
__device__ __forceinline__ int xcd_remap(int p, int nb) {
    const int q = nb / kNumXcd, rem = nb % kNumXcd, r = p % kNumXcd;
    return r * q + (r < rem ? r : rem) + p / kNumXcd;
}

__device__ __forceinline__ Ctx bubble_ctx(Foam *foams, const int2 *bt) {
    const int2 e = bt[blockIdx.x];
    Foam *F = foams + e.x;
    return Ctx{F, xcd_remap(e.y, F->nblk)};
}

__device__ __forceinline__ Ctx cell_ctx(Foam *foams, const int2 *bt) {
    const int2 e = bt[blockIdx.x];
    return Ctx{foams + e.x, e.y};
}

// ---- deterministic reductions (fixed butterfly, every lane identical) ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[kWaves];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) t += sh[k];
    return t;
}

__device__ __forceinline__ double block_max(double v) {
    __shared__ double sh[kWaves];
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = sh[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) t = fmax(t, sh[k]);
    return t;
}

__device__ __forceinline__ int block_sum_i(int v) {
    __shared__ int sh[kWaves];
    v = wave_sum_i(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) t += sh[k];
    return t;
}

__device__ __forceinline__ int block_max_i(int v) {
    __shared__ int sh[kWaves];
    v = wave_max_i(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = sh[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) t = max(t, sh[k]);
    return t;
}

// Exclusive scan of one int per thread over the block, in thread order.
__device__ __forceinline__ int block_excl_scan(int v, int &total) {
    __shared__ int sw[kWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    __syncthreads();
    if (lane == 63) sw[w] = x;
    __syncthreads();
    int wo = 0, t = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
        if (k < w) wo += sw[k];
        t += sw[k];
    }
    total = t;
    return wo + x - v;
}

// ---- geometry (bit-identical to the FP64 restatement) --------------------

// Minimum-image difference a - b: reference src/kernels.cu:1094-1114.
template <int D>
__device__ __forceinline__ void wrapped_diff(const Consts &c, double x1, double y1, double z1,
                                             double x2, double y2, double z2, double d[3]) {
    double d1[3] = {x1 - x2, y1 - y2, D == 3 ? z1 - z2 : 0.0};
    d[0] = d1[0];
    d[1] = d1[1];
    d[2] = d1[2];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double a = d1[k] < 0 ? -d1[k] : d1[k];
        const double t = c.interval[k] - a;
        if (!c.wall[k] && t * t < d1[k] * d1[k]) d[k] = t * (d1[k] < 0 ? 1.0 : -1.0);
    }
}

// Squared length, x, y, z summed in order: reference src/vec.h:41-49.
__device__ __forceinline__ double sq_len(const double d[3]) {
    double t = 0;
    t += d[0] * d[0];
    t += d[1] * d[1];
    t += d[2] * d[2];
    return t;
}

// Cell grid from the radius bound: reference src/simulator.cu:69-75,
// clamped to >= 1 per axis and shrunk (largest axis first) to the capacity.
__device__ __forceinline__ void cell_dims(const Consts &c, double max_r, int cap, int cd[3]) {
    const double edge = 2 * (max_r + c.skin);
    for (int k = 0; k < 3; ++k) {
        double q = k < c.dim ? floor(c.interval[k] / edge) : 1.0;
        if (!(q >= 1.0)) q = 1.0;
        if (q > 1.0e6) q = 1.0e6;
        cd[k] = (int)q;
    }
    if (2 == c.dim) cd[2] = 1;
    while ((long long)cd[0] * cd[1] * cd[2] > (long long)cap) {
        int big = 0;
        for (int k = 1; k < 3; ++k)
            if (cd[k] > cd[big]) big = k;
        cd[big] -= 1;
    }
}

// Newton cube root from correctly rounded operations only (frexp / ldexp
// are exact), so the volume fix rounds identically on host and device.
__device__ __forceinline__ double det_cbrt(double a) {
    if (!(a > 0.0) || isinf(a)) return a;
    int e = 0;
    double m = frexp(a, &e);
    int r = e % 3;
    if (r < 0) r += 3;
    m = ldexp(m, r);
    e -= r;
    double y = 1.0;
    for (int it = 0; it < 8; ++it) {
        const double y2 = y * y;
        const double f = y2 * y - m;
        const double g = 3.0 * y2;
        y = y - f / g;
    }
    return ldexp(y, e / 3);
}

}  // namespace cubble_amd
