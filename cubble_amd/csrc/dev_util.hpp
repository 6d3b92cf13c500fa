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

// Streaming store of a result the next kernel reads: marked nontemporal, so
// it does not stay dirty in this XCD's L2 (the kernel boundary writes dirty
// lines back anyway; L2s of different XCDs are not coherent).
template <typename T>
__device__ __forceinline__ void stream_store(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}

// A per-bubble buffer stored as double in FP64 mode and as float in mixed
// precision (velocities, saved positions, error, overlap), as its element type.
template <typename V>
__device__ __forceinline__ V *field(void *p) {
    return static_cast<V *>(p);
}

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

// Several block reductions in one LDS round: each wave reduces its values
// with the butterflies above, lane 0 stores them, and after one barrier
// thread 0 combines the waves in wave order, as block_sum, block_max and
// block_sum_i do (the same bits).  Bit k of `dmax` makes d[k] a maximum,
// else a sum; `iv` is summed.  The results are defined in thread 0 only.
// The leading barrier guards the LDS against a previous call.
template <int ND>
__device__ __forceinline__ void block_reduce_t0(double (&d)[ND], unsigned dmax, int &iv) {
    __shared__ double sd[kWaves][ND];
    __shared__ int si[kWaves];
#pragma unroll
    for (int k = 0; k < ND; ++k) d[k] = (dmax >> k) & 1u ? wave_max(d[k]) : wave_sum(d[k]);
    iv = wave_sum_i(iv);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < ND; ++k) sd[threadIdx.x >> 6][k] = d[k];
        si[threadIdx.x >> 6] = iv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            double t = (dmax >> k) & 1u ? sd[0][k] : 0.0 + sd[0][k];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) t = (dmax >> k) & 1u ? fmax(t, sd[w][k]) : t + sd[w][k];
            d[k] = t;
        }
        int t = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += si[w];
        iv = t;
    }
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
#pragma clang fp contract(off)
    double d1[3] = {x1 - x2, y1 - y2, D == 3 ? z1 - z2 : 0.0};
    d[0] = d1[0];
    d[1] = d1[1];
    d[2] = d1[2];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double a = d1[k] < 0 ? -d1[k] : d1[k];
        const double t = c.interval[k] - a;
        // selects, not branches: the same operations and rounding either way
        const double w = d1[k] < 0 ? t : -t;
        d[k] = (!c.wall[k] && t * t < d1[k] * d1[k]) ? w : d1[k];
    }
}

// Lane l's double, broadcast to the wave (l must be wave-uniform).
__device__ __forceinline__ double readlane_d(double v, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Neighbour gather of bubble i over its list (column-major, nbr[q * cap + i])
// in chunks of kGather: the chunk's list entries are loaded, then all their
// records, then `body(j, record)` runs for each in ascending list order.  A
// lane's loads of one chunk are independent, so their latencies overlap
// instead of adding up entry by entry.  Past the end of the list a lane
// re-reads its own record and skips the body.
constexpr int kGather = 4;

// The buffers live in device global memory; pointers read from the foam
// header are generic, so loads through them compile to flat loads, whose
// completion waits also cover LDS traffic.  gmem() restates the address
// space so they compile to global loads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}

// For data a kernel only reads: the constant address space lets wave-uniform
// loads go through the scalar cache into SGPRs.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *cmem(const T *p) {
    return (const __attribute__((address_space(4))) T *)p;
}

// One neighbour record: the 32-byte FP64 record as two 16-byte loads, the
// 16-byte fixed-point record as one.
__device__ __forceinline__ D4 load_rec(const __attribute__((address_space(1))) D4 *p) {
    return D4{p->x, p->y, p->z, p->r};
}

__device__ __forceinline__ Q4 load_rec(const __attribute__((address_space(1))) Q4 *p) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i v = *(const __attribute__((address_space(1))) v4i *)p;
    return Q4{v.x, v.y, v.z, __int_as_float(v.w)};
}

template <typename Rec, typename Body>
__device__ __forceinline__ void for_neighbours(const int *__restrict__ nbr_,
                                               const Rec *__restrict__ p4_, size_t cap, int i,
                                               int m_nn, Body &&body) {
    const auto nbr = gmem(nbr_);
    const auto p4 = gmem(p4_);
    if (m_nn <= 0) return;
    // list entries of batch q0 (past the end: the last entry, in bounds; the
    // entries past the end are masked off)
    auto load_ids = [&](int q0, int js[kGather]) {
#pragma unroll
        for (int u = 0; u < kGather; ++u) {
            const int q = q0 + u < m_nn ? q0 + u : m_nn - 1;
            const int j = nbr[(size_t)q * cap + i];
            js[u] = q0 + u < m_nn ? j : -1;
        }
    };
    // software pipelined: the next batch's list entries are loaded while the
    // current batch's records are in flight
    int js[kGather];
    load_ids(0, js);
    for (int q0 = 0; q0 < m_nn; q0 += kGather) {
        Rec pj[kGather];
#pragma unroll
        for (int u = 0; u < kGather; ++u) {
            const int j = js[u] >= 0 ? js[u] : i;
            pj[u] = load_rec(p4 + j);
        }
        int cur[kGather];
#pragma unroll
        for (int u = 0; u < kGather; ++u) cur[u] = js[u];
        if (q0 + kGather < m_nn) load_ids(q0 + kGather, js);
#pragma unroll
        for (int u = 0; u < kGather; ++u)
            if (cur[u] >= 0) body(cur[u], pj[u]);
    }
}

// ---- fixed-point positions of the mixed-precision neighbour record ------
// Code of coordinate x on axis (lb, L, wall): periodic axes take
// round((x - lb) 2^32 / L) modulo 2^32, as a signed 32-bit value offset by
// 2^31 (the offset cancels in differences); positions a little outside the
// box wrap like the box does.  Wall axes take round((x - lb) 2^30 / L),
// clamped to the int32 range.
__device__ __forceinline__ int32_t fixed_code(double x, double lb, double L, bool wall) {
    if (wall) {
        double t = rint((x - lb) * (0x1p30 / L));
        t = fmin(fmax(t, -2147483647.0), 2147483647.0);
        return (int32_t)t;
    }
    double t = (x - lb) * (0x1p32 / L);
    t -= 0x1p32 * floor(t * 0x1p-32);  // [0, 2^32)
    t = rint(t - 0x1p31);               // [-2^31, 2^31]
    if (t >= 0x1p31) t -= 0x1p32;
    return (int32_t)t;
}

// Length of one code unit on an axis, in FP32.
__device__ __forceinline__ float fixed_step(double L, bool wall) {
    return (float)(L * (wall ? 0x1p-30 : 0x1p-32));
}

// Minimum-image (periodic) or plain (wall) difference a - b of two codes, in
// code units: the wrapping 32-bit difference.
__device__ __forceinline__ int fixed_diff(int32_t a, int32_t b) {
    return (int)((uint32_t)a - (uint32_t)b);
}

// FP64 reciprocal and reciprocal square root: the hardware approximations
// (v_rcp_f64, v_rsq_f64) refined by one Newton step with FMA, within a few
// ulp of the correctly rounded 1 / x and 1 / sqrt(x) -- far inside the FP64
// path's 1e-12 tolerance, at a fraction of the cost of IEEE division and
// square root.  x > 0 and finite.
__device__ __forceinline__ double rcp_nr(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(fma(-x, r, 1.0), r, r);
}

__device__ __forceinline__ double rsq_nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    return fma(fma(-hx * y, y, 0.5), y, y);  // y (1.5 - x y^2 / 2)
}

// AB2 predicted position and radius of bubble i: reference
// src/kernels.cu:48-50.  Shared by k_predict and, in mixed precision (where
// no FP64 record is stored), by k_correct and the pair kernel's flow test, so
// all round identically (no contraction, whatever the file's setting).
template <int D, bool GAS, typename V>
__device__ __forceinline__ D4 predicted(const Roles &R, double h, int i) {
#pragma clang fp contract(off)
    double p[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < D; ++q)
        p[q] = R.x[q][i] + h * (3.0 * (double)field<V>(R.v[q])[i] - (double)field<V>(R.vo[q])[i]);
    const double rp =
        GAS ? R.x[3][i] + h * (3.0 * (double)field<V>(R.v[3])[i] - (double)field<V>(R.vo[3])[i])
            : R.x[3][i];
    return D4{p[0], p[1], p[2], rp};
}

// Imposed flow (reference addFlowVelocity, src/kernels.cu:497-544).  Inside
// the flow box -- per axis strictly between the faces, or within the
// bubble's radius of either face -- the bubble takes the imposed velocity;
// elsewhere the mean velocity of its neighbours: their velocity sum `fv`
// times 1 / deg, deg being its list length at the last search.  p is the
// bubble's predicted position and radius (FP64); the term is added to the
// pair sum `acc`, before the walls.
template <int D, typename T>
__device__ __forceinline__ void add_flow(const Consts &c, const D4 &p, int deg, const T fv[3],
                                         T acc[3]) {
#pragma clang fp contract(off)
    const double r2 = p.r * p.r;
    const double x[3] = {p.x, p.y, p.z};
    bool inside = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const double lo = c.flow_lbb[a] - x[a], hi = c.flow_tfr[a] - x[a];
        inside = inside && ((x[a] < c.flow_tfr[a] && x[a] > c.flow_lbb[a]) || lo * lo <= r2 ||
                            hi * hi <= r2);
    }
    const double mult = deg > 0 ? 1.0 / (double)deg : 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) acc[a] += inside ? (T)c.flow_vel[a] : (T)(mult * (double)fv[a]);
}

// Squared length, x, y, z summed in order: reference src/vec.h:41-49.
__device__ __forceinline__ double sq_len(const double d[3]) {
#pragma clang fp contract(off)
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
