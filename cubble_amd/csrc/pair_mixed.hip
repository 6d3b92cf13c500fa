// pair_mixed.hip -- mixed-precision pairwise kernel.
//
// Same gather structure as the FP64 pair kernel (ascending neighbour order,
// deterministic), over a 16-byte neighbour record written by the predictor:
// the predicted position in 32-bit fixed point over the box and the radius in
// FP32 (one 16-byte load per neighbour instead of two).  The relative
// coordinate d = wrap(x_i - x_j) is the exact integer difference of two codes
// (the wrap is the integer wrap-around on periodic axes), rounded once to
// FP32; a code unit is L / 2^32 (L / 2^30 on wall axes), so quantisation adds
// at most L / 2^32 per component.  The Durian velocity, overlap area and gas
// flux are evaluated in FP32 with the hardware's one-instruction reciprocal
// square root, reciprocal and square root (about 1 ulp; the correctly
// rounded variants compile to division sequences), and accumulated per
// bubble in FP32, and stored as FP32 (the
// velocity buffers of mixed precision).  Positions, radii and the corrector's
// arithmetic stay FP64.  The overlap area uses the product
// form
//     a = (R^2 - |d|^2) (|d|^2 - (r1 - r2)^2) / (4 |d|^2),
// algebraically equal to the reference's r2^2 - ((r2^2 - r1^2 + |d|^2)/2)^2/|d|^2
// (src/kernels.cu:240-249) but without its cancellation near tangency.
//
// The reference's overlap area jumps where one centre enters the other
// bubble (|d| < max(r1, r2): the area becomes min(r1, r2)^2, :238-239), and
// the imposed flow counts a larger neighbour only past the overlap test
// (:281-307).  FP32 distances would decide those two branch points
// differently from FP64 for pairs that close to them, and one such pair moves
// its bubbles' overlap and dR/dt by O(1).  Near a threshold |d| = t the FP32
// value of |d|^2 - t^2 is off by at most 2 sqrt(3) s |d| from the code
// quantisation (s the largest code unit) plus a few FP32 roundings of |d|^2
// (~5e-7 relative).  The band is four times that: |d|^2 - t^2 within
// 8 sqrt(3) s t + 2^-18 t^2.
//
// Two more kinds of pair need more than FP32 values from codes.  Code
// quantisation is an absolute error (up to sqrt(D) s in d), while the
// velocity term f0 (d / |d| - d / R) needs d to a relative error: for nearly
// coincident centres the direction d / |d| is off by up to sqrt(D) s / |d|,
// so pairs closer than 2^18 sqrt(D) s count (such pairs exist in freshly
// seeded foams and part within a few steps).  In 2D the overlap term
// 2 sqrt(a2) has a2 proportional to R^2 - |d|^2 (R = r1 + r2), so near
// tangency an error e in that difference moves the term by up to
// e / (2 sqrt|R^2 - |d|^2|); e is at most 2 sqrt(2) s R + 11 2^-24 R^2
// (quantisation, and the roundings of the FP32 radii and of |d|^2), so
// 2D pairs within (sqrt(2) 2^15 s + 11 2^-10 R)^2 of tangency count, and
// every other term stays within 2^-15 R.  (In 3D the term is linear in the
// difference.)
//
// A bubble that met such a pair is flagged (the sign of its stored overlap
// sum) and its workgroup marked; a second kernel sums each flagged bubble
// again with one wave, taking those pairs' d, |d|^2, R^2 - |d|^2 and both
// branch decisions from the FP64 predictions, exactly as the FP64 path
// evaluates them (rounded once to FP32 where the formula continues in FP32),
// in parallel over the marked workgroups of every foam.
#include "dev_util.hpp"
#include "launch.hpp"

namespace cubble_amd {

// Relative part of the band around a branch threshold that is decided in
// FP64; its absolute part is kBandSteps code units per unit of |d|.
constexpr float kBranchBand = 0x1p-18f;
constexpr float kBandSteps = 13.856406f;  // 8 sqrt(3)
// Pairs closer than kCloseSteps sqrt(D) code units take their relative
// coordinate from the FP64 predictions.
constexpr float kCloseSteps = 0x1p18f;
// 2D pairs with |(r1 + r2)^2 - |d|^2| <= (kTangentS s + kTangentR (r1 + r2))^2
// take it from the FP64 predictions.
constexpr float kTangentS = 46340.950f;  // sqrt(2) 2^15
constexpr float kTangentR = 0x1.6p-7f;   // 11 2^-10

// Pair (i, j) as the FP64 path evaluates it, from the FP64 predictions
// (recomputed as k_correct does; pi is bubble i's): the relative coordinate
// x_i - x_j (minimum image), |d|^2 and (r1 + r2)^2 - |d|^2, each rounded once
// to FP32, and the FP64 path's two branch decisions: the pair overlaps
// ((r1 + r2)^2 >= |d|^2), and a centre lies inside the other bubble
// (|d|^2 < r1^2 or |d|^2 < r2^2), r1 the radius of the smaller index.
struct ExactPair {
    float dd[3], m, over;
    bool overlap, inside;
};

template <int D, bool GAS>
__device__ __forceinline__ ExactPair exact_pair(const Consts &c, const Roles &R, double h,
                                                const D4 &pi, int i, int j) {
#pragma clang fp contract(off)
    const D4 pj = predicted<D, GAS, float>(R, h, j);
    double d[3];
    wrapped_diff<D>(c, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z, d);
    const double m = sq_len(d);
    const double r1 = i < j ? pi.r : pj.r;
    const double r2 = i < j ? pj.r : pi.r;
    const double radii = r1 + r2;
    const double r1sq = r1 * r1, r2sq = r2 * r2;
    ExactPair e;
    e.dd[0] = (float)d[0];
    e.dd[1] = (float)d[1];
    e.dd[2] = D == 3 ? (float)d[2] : 0.f;
    e.m = (float)m;
    e.over = (float)(radii * radii - m);
    e.overlap = radii * radii >= m;
    e.inside = m < r1sq || m < r2sq;
    return e;
}

struct MixedOut {
    float acc[3], fv[3], ovl, dr;
};

// The terms of bubble i's pairs, one call per neighbour j.  EXACT = false
// evaluates the pair in FP32 from code differences; EXACT = true takes the
// values and decisions of a pair that needs them from the FP64 predictions
// (pi, bubble i's), with everything else the same.  A call adds the pair's
// terms to o and returns whether the pair needs FP64 values: nearly
// coincident centres, near-tangent in 2D, or near a branch point.
template <int D, bool GAS, bool FLOW, bool EXACT>
struct MixedPair {
    const Consts &c;
    const Roles &R;
    const float *step;
    double h;
    int i;
    Q4 qi;
    D4 pi;
    float f0, ri_f, inv_ri, band_abs, close2, tangent_s, inside_k0, inside_k1;

    __device__ __forceinline__ MixedPair(const Consts &c_, const Roles &R_, double h_,
                                         const float step_[3], int i_, const Q4 &qi_)
        : c(c_), R(R_), step(step_), h(h_), i(i_), qi(qi_), pi{} {
        f0 = (float)c.f0;
        ri_f = qi.r;
        inv_ri = __builtin_amdgcn_rcpf(ri_f);
        // absolute part of the threshold band: the largest code unit
        const float s_max = fmaxf(fmaxf(step[0], step[1]), D == 3 ? step[2] : 0.f);
        band_abs = kBandSteps * s_max;
        // distance below which a pair's d comes from the FP64 predictions
        const float close = kCloseSteps * (D == 3 ? 1.7320508f : 1.4142135f) * s_max;
        close2 = close * close;
        // absolute part of the 2D tangency band
        tangent_s = kTangentS * s_max;
        // the band around the inside threshold big = max(r1, r2)^2, at least
        // band_abs max(r1, r2) + kBranchBand big
        inside_k0 = 0.5f * band_abs;
        inside_k1 = 0.5f * band_abs + kBranchBand;
        if (EXACT) pi = predicted<D, GAS, float>(R, h, i);
    }

    __device__ __forceinline__ bool operator()(int j, const Q4 &qj, MixedOut &o) const {
        const float pi_f = (float)kPi;
        const float rj_f = qj.r;
        // code differences are exact integers; one rounding to FP32 each
        float dd[3] = {(float)fixed_diff(qi.x, qj.x) * step[0],
                       (float)fixed_diff(qi.y, qj.y) * step[1],
                       D == 3 ? (float)fixed_diff(qi.z, qj.z) * step[2] : 0.f};
        float m = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
        const float rad = ri_f + rj_f;
        float over = rad * rad - m;
        // a centre inside the other bubble: |d|^2 < r1^2 or r2^2, that is
        // below the larger square; the product form of the area is symmetric
        // in the radii, so their order does not matter
        const float sq_i = ri_f * ri_f, sq_j = rj_f * rj_f;
        const float big = fmaxf(sq_i, sq_j);
        // pairs the FP32 values do not settle (each test evaluated: a wave
        // rarely skips one, so no branches)
        bool special = m < close2;
        if (FLOW)
            special |= j > i && fabsf(over) <= band_abs * rad + kBranchBand * (rad * rad);
        if (GAS && D == 2) {
            const float t = fmaf(kTangentR, rad, tangent_s);
            special |= fabsf(over) <= t * t;
        }
        // (the inside band uses max(r1, r2) <= (big + 1) / 2: one FMA)
        if (GAS) special |= fabsf(m - big) <= fmaf(big, inside_k1, inside_k0);
        bool overlap = over >= 0.f;
        bool inside = m < big;
        if (EXACT && special) {
            const ExactPair e = exact_pair<D, GAS>(c, R, h, pi, i, j);
#pragma unroll
            for (int a = 0; a < 3; ++a) dd[a] = e.dd[a];
            m = e.m;
            over = e.over;
            overlap = e.overlap;
            inside = e.inside;
        }
        if (FLOW && (j < i || overlap)) {
#pragma unroll
            for (int a = 0; a < D; ++a) o.fv[a] += field<float>(R.v[a])[j];
        }
        if (overlap) {
            const float rs = __builtin_amdgcn_rsqf(m);
            const float s = rs - __builtin_amdgcn_rcpf(rad);
#pragma unroll
            for (int a = 0; a < D; ++a) o.acc[a] += dd[a] * (f0 * s);
            if (GAS) {
                float a2;
                if (inside) {
                    a2 = fminf(sq_i, sq_j);
                } else {
                    const float dl = ri_f - rj_f;
                    a2 = fabsf(over * (m - dl * dl) * (0.25f * rs * rs));
                }
                const float oa = 3 == D ? a2 * pi_f : 2.0f * __builtin_amdgcn_sqrtf(a2);
                o.ovl += oa;
                o.dr += oa * (__builtin_amdgcn_rcpf(rj_f) - inv_ri);
            }
        }
        return special;
    }
};

__device__ __forceinline__ void zero_out(MixedOut &o) {
#pragma unroll
    for (int a = 0; a < 3; ++a) o.acc[a] = o.fv[a] = 0.f;
    o.ovl = o.dr = 0.f;
}

// Bubble i's terms stored: velocities, dR/dt and overlap, the overlap's sign
// bit set when `flag` (the overlap sum is never negative: the flag marks a
// bubble that met a pair needing FP64 values, for the second kernel, which
// stores it plain).  FLOW: the imposed-flow term is added first (the
// neighbour-velocity sum in FP32, the flow-box test on the bubble's FP64
// prediction, recomputed here as the corrector does).
template <int D, bool GAS, bool FLOW>
__device__ __forceinline__ void store_mixed(const Consts &c, const Roles &R, const Scratch &S,
                                            double h, int i, MixedOut &o, bool flag) {
    if (FLOW) add_flow<D>(c, predicted<D, GAS, float>(R, h, i), S.deg[i], o.fv, o.acc);
#pragma unroll
    for (int a = 0; a < D; ++a) stream_store(field<float>(R.vp[a]) + i, o.acc[a]);
    stream_store(field<float>(R.vp[3]) + i, o.dr);
    stream_store(field<float>(S.ov) + i,
                 flag ? __uint_as_float(__float_as_uint(o.ovl) | 0x80000000u) : o.ovl);
}

// The bubbles of bubble-space block lb (thread t takes base + k kBlock + t,
// k < ipt): each bubble's pair sums over its list, ascending, in FP32, its
// terms stored, and the block's overlap sums, per thread in k order.
template <int D, bool GAS, bool FLOW>
__device__ __forceinline__ void mixed_block(const Consts &c, const Roles &R, const Scratch &S,
                                            size_t cap, int base, int ipt, int n, double h,
                                            double &soa, double &soar, bool &near_any) {
    float step[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) step[a] = fixed_step(c.interval[a], c.wall[a] != 0);
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        const Q4 qi = load_rec(gmem(S.q4) + i);
        const MixedPair<D, GAS, FLOW, false> pair(c, R, h, step, i, qi);
        MixedOut o;
        zero_out(o);
        bool near = false;
        for_neighbours(S.nbr, S.q4, cap, i, S.nn[i],
                       [&](int j, const Q4 &qj) { near = pair(j, qj, o) || near; });
        store_mixed<D, GAS, FLOW>(c, R, S, h, i, o, near);
        near_any = near_any || near;
        soa += (double)o.ovl;
        soar += (double)o.ovl * (1.0 / (double)qi.r);
    }
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Flagged bubble i (wave-uniform) summed again by the whole wave, the values
// of the pairs that need them taken from the FP64 predictions: lane l takes
// list entries l, l + 64, ..., and the lanes' sums are added in a fixed
// butterfly order (every lane ends with the same sums, independent of
// timing).  Lane 0 stores the terms; returns the overlap sum.
template <int D, bool GAS, bool FLOW>
__device__ __forceinline__ float exact_bubble_wave(const Consts &c, const Roles &R,
                                                   const Scratch &S, size_t cap, double h,
                                                   const float step[3], int i) {
    const Q4 qi = load_rec(gmem(S.q4) + i);
    const MixedPair<D, GAS, FLOW, true> pair(c, R, h, step, i, qi);
    MixedOut o;
    zero_out(o);
    const int m_nn = S.nn[i];
    for (int q = (int)(threadIdx.x & 63); q < m_nn; q += 64) {
        const int j = gmem(S.nbr)[(size_t)q * cap + i];
        pair(j, load_rec(gmem(S.q4) + j), o);
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        o.acc[a] = wave_sum_f(o.acc[a]);
        if (FLOW) o.fv[a] = wave_sum_f(o.fv[a]);
    }
    if (GAS) {
        o.ovl = wave_sum_f(o.ovl);
        o.dr = wave_sum_f(o.dr);
    }
    if (0 == (threadIdx.x & 63)) store_mixed<D, GAS, FLOW>(c, R, S, h, i, o, false);
    return o.ovl;
}

// The flagged bubbles of block lb, one wave per flagged bubble, and the
// block's overlap sums formed again: the other bubbles' stored overlaps and
// radii, the flagged ones' new overlaps, per thread in k order as the main
// pass adds them.
template <int D, bool GAS, bool FLOW>
__device__ __forceinline__ void fix_block(const Consts &c, const Roles &R, const Scratch &S,
                                          size_t cap, int base, int ipt, int n, double h,
                                          double &soa, double &soar) {
    __shared__ int s_cnt;
    __shared__ int s_list[kBlock];
    __shared__ float s_ovl[kBlock];
    float step[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) step[a] = fixed_step(c.interval[a], c.wall[a] != 0);
    const int wave = (int)(threadIdx.x >> 6);
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        float ovl = 0.f, ri = 1.f;
        int slot = -1;
        if (i < n) {
            ovl = field<float>(S.ov)[i];
            ri = load_rec(gmem(S.q4) + i).r;
            if (__float_as_uint(ovl) >> 31) {
                slot = atomicAdd(&s_cnt, 1);
                s_list[slot] = i;
            }
        }
        __syncthreads();
        const int cnt = s_cnt;
        for (int f = wave; f < cnt; f += kBlock / 64) {
            const float o = exact_bubble_wave<D, GAS, FLOW>(c, R, S, cap, h, step, s_list[f]);
            if (0 == (threadIdx.x & 63)) s_ovl[f] = o;
        }
        __syncthreads();
        if (slot >= 0) ovl = s_ovl[slot];
        if (i < n) {
            soa += (double)ovl;
            soar += (double)ovl * (1.0 / (double)ri);
        }
    }
}

// Foam constants and the (dimension, gas) variant of block lb; FIX selects
// the second kernel's block pass.
template <bool FLOW, bool FIX>
__device__ __forceinline__ void mixed_block_dispatch(const Foam *F, int lb, uint32_t flags,
                                                     double &soa, double &soar, bool &near_any) {
    const int n = F->d.count, ipt = F->ipt;
    const Consts c = F->c;
    const Roles R = F->r;
    const Scratch S = F->s;
    const size_t cap = (size_t)F->cap;
    const int base = lb * kBlock * ipt;
    const double h = 0.5 * F->d.dt;
    const bool gas = (flags & kGas) != 0;
    if (3 == c.dim) {
        if (gas) {
            if (FIX) fix_block<3, true, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar);
            else mixed_block<3, true, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar, near_any);
        } else {
            if (FIX) fix_block<3, false, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar);
            else mixed_block<3, false, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar, near_any);
        }
    } else {
        if (gas) {
            if (FIX) fix_block<2, true, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar);
            else mixed_block<2, true, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar, near_any);
        } else {
            if (FIX) fix_block<2, false, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar);
            else mixed_block<2, false, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar, near_any);
        }
    }
}

// One kernel per FLOW value (the flow variant needs more registers; the
// plain one is held to 64 VGPRs, eight waves per SIMD).  Every branch is
// decided in FP32; a bubble that met a pair that needs FP64 values carries a
// flag in the sign of its stored overlap sum, and its wave sets a bit in the
// block's mark (block partial slot 2).
template <bool FLOW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(FLOW ? 1 : 8, 8))) void
k_pair_mixed(Foam *foams, const int2 *bt, uint32_t flags) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (F->d.halt) return;
    double soa = 0.0, soar = 0.0;
    bool near = false;
    mixed_block_dispatch<FLOW, false>(F, cx.lb, flags, soa, soar, near);
    const bool wave_near = __builtin_amdgcn_ballot_w64(near) != 0ull;
    const int wave_bit = 0 == (threadIdx.x & 63) && wave_near ? 1 << (threadIdx.x >> 6) : 0;
    double sums[2] = {soa, soar};
    int mark = wave_bit;
    block_reduce_t0(sums, 0u, mark);
    if (threadIdx.x == 0) {
        F->s.part[cx.lb * kPartD + 2] = sums[0];
        F->s.part[cx.lb * kPartD + 3] = sums[1];
        F->s.part_i[cx.lb * kPartI + 2] = mark;
    }
}

// The flagged bubbles of every foam, in parallel, one wave each: summed again
// with the values and decisions of the pairs that need them taken from the
// FP64 predictions, and their blocks' overlap partials formed again.  An
// unmarked workgroup returns at once.
template <bool FLOW>
__global__ __launch_bounds__(kBlock) void k_fix_mixed(Foam *foams, const int2 *bt,
                                                      uint32_t flags) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (F->d.halt) return;
    const int mark = F->s.part_i[cx.lb * kPartI + 2];
    if (!mark) return;
    double soa = 0.0, soar = 0.0;
    bool near = false;
    mixed_block_dispatch<FLOW, true>(F, cx.lb, flags, soa, soar, near);
    double sums[2] = {soa, soar};
    int none = 0;
    block_reduce_t0(sums, 0u, none);
    if (threadIdx.x == 0) {
        F->s.part[cx.lb * kPartD + 2] = sums[0];
        F->s.part[cx.lb * kPartD + 3] = sums[1];
    }
}

void launch_pair_mixed(const Grid &g, uint32_t flags, bool fix) {
    if (flags & kFlow) {
        hipLaunchKernelGGL(k_pair_mixed<true>, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams,
                           g.bt_bub, flags);
        if (fix)
            hipLaunchKernelGGL(k_fix_mixed<true>, dim3(g.nb_bub), dim3(kBlock), 0, g.stream,
                               g.foams, g.bt_bub, flags);
    } else {
        hipLaunchKernelGGL(k_pair_mixed<false>, dim3(g.nb_bub), dim3(kBlock), 0, g.stream,
                           g.foams, g.bt_bub, flags);
        if (fix)
            hipLaunchKernelGGL(k_fix_mixed<false>, dim3(g.nb_bub), dim3(kBlock), 0, g.stream,
                               g.foams, g.bt_bub, flags);
    }
}

}  // namespace cubble_amd
