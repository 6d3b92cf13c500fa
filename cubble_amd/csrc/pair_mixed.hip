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
// A wave whose bubbles met such a pair marks itself, and a second kernel sums
// the marked waves' bubbles again taking those pairs' d, |d|^2, R^2 - |d|^2
// and both branch decisions from the FP64 predictions, exactly as the FP64
// path evaluates them (rounded once to FP32 where the formula continues in
// FP32), in parallel over the marked waves of every foam.
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

// One bubble's pair sums over its list, ascending.  EXACT = false evaluates
// every pair in FP32 from code differences and reports whether a pair needs
// FP64 values (close, near-tangent in 2D, or near a branch point); EXACT =
// true (run only for such bubbles) takes those pairs' values and decisions
// from the FP64 predictions, with everything else the same.
template <int D, bool GAS, bool FLOW, bool EXACT>
__device__ __forceinline__ bool pair_bubble_mixed(const Consts &c, const Roles &R,
                                                  const Q4 *__restrict__ q4,
                                                  const int *__restrict__ nbr, size_t cap,
                                                  double h, const float step[3], int i,
                                                  const Q4 &qi, int m_nn, MixedOut &o) {
    const float f0 = (float)c.f0;
    const float pi_f = (float)kPi;
    const float ri_f = qi.r;
    // absolute part of the threshold band: the largest code unit
    const float s_max = fmaxf(fmaxf(step[0], step[1]), D == 3 ? step[2] : 0.f);
    const float band_abs = kBandSteps * s_max;
    // distance below which a pair's d comes from the FP64 predictions
    const float close = kCloseSteps * (D == 3 ? 1.7320508f : 1.4142135f) * s_max;
    const float close2 = close * close;
    // absolute part of the 2D tangency band
    const float tangent_s = kTangentS * s_max;
    // the band around the inside threshold big = max(r1, r2)^2, at least
    // band_abs max(r1, r2) + kBranchBand big
    const float inside_k0 = 0.5f * band_abs, inside_k1 = 0.5f * band_abs + kBranchBand;
    const float inv_ri = __builtin_amdgcn_rcpf(ri_f);
    D4 pi{};
    if (EXACT) pi = predicted<D, GAS, float>(R, h, i);
    bool near = false;
    for (int a = 0; a < 3; ++a) o.acc[a] = o.fv[a] = 0.f;
    o.ovl = o.dr = 0.f;
    auto body = [&](int j, const Q4 &qj) {
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
        // pairs the FP32 values do not settle: nearly coincident centres,
        // near-tangent pairs in 2D, and pairs near a branch point
        bool special = m < close2;
        if (FLOW)
            special = special || (j > i && fabsf(over) <= band_abs * rad + kBranchBand * (rad * rad));
        if (GAS && D == 2) {
            const float t = fmaf(kTangentR, rad, tangent_s);
            special = special || fabsf(over) <= t * t;
        }
        // (the inside band uses max(r1, r2) <= (big + 1) / 2: one FMA)
        if (GAS) special = special || fabsf(m - big) <= fmaf(big, inside_k1, inside_k0);
        bool overlap = over >= 0.f;
        bool inside = m < big;
        if (special) {
            near = true;
            if (EXACT) {
                const ExactPair e = exact_pair<D, GAS>(c, R, h, pi, i, j);
#pragma unroll
                for (int a = 0; a < 3; ++a) dd[a] = e.dd[a];
                m = e.m;
                over = e.over;
                overlap = e.overlap;
                inside = e.inside;
            }
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
                    a2 = over * (m - dl * dl) * (0.25f * rs * rs);
                    a2 = a2 < 0.f ? -a2 : a2;
                }
                const float oa = 3 == D ? a2 * pi_f : 2.0f * __builtin_amdgcn_sqrtf(a2);
                o.ovl += oa;
                o.dr += oa * (__builtin_amdgcn_rcpf(rj_f) - inv_ri);
            }
        }
    };
    // (the EXACT pass runs in its own kernel, so its FP64 work does not add
    // to the registers of the main pass)
    for_neighbours(nbr, q4, cap, i, m_nn, body);
    return near;
}

// One bubble's pair terms, stored (velocities, dR/dt, overlap); returns its
// overlap sum and radius for the block partials, and whether a pair came
// a pair that needs FP64 values.  FLOW: also the imposed-flow term
// (the neighbour-velocity sum in FP32, the flow-box test on the bubble's FP64
// prediction, recomputed here as the corrector does); same neighbour
// selection as the FP64 kernel.
template <int D, bool GAS, bool FLOW, bool EXACT>
__device__ __forceinline__ float mixed_bubble(const Consts &c, const Roles &R, const Scratch &S,
                                              size_t cap, double h, const float step[3], int i,
                                              float &ri, bool &near) {
    const Q4 qi = load_rec(gmem(S.q4) + i);
    MixedOut o;
    near = pair_bubble_mixed<D, GAS, FLOW, EXACT>(c, R, S.q4, S.nbr, cap, h, step, i, qi,
                                                  S.nn[i], o);
    if (FLOW) add_flow<D>(c, predicted<D, GAS, float>(R, h, i), S.deg[i], o.fv, o.acc);
#pragma unroll
    for (int a = 0; a < D; ++a) stream_store(field<float>(R.vp[a]) + i, o.acc[a]);
    stream_store(field<float>(R.vp[3]) + i, o.dr);
    // the overlap sum is never negative: in the main pass its sign bit flags a
    // bubble that met a pair needing FP64 values, for the second kernel, whose
    // pass stores it plain
    stream_store(field<float>(S.ov) + i,
                 !EXACT && near ? __uint_as_float(__float_as_uint(o.ovl) | 0x80000000u)
                                : o.ovl);
    ri = qi.r;
    return o.ovl;
}

// The bubbles of bubble-space block lb (thread t takes base + k kBlock + t,
// k < ipt): pair terms and the block's overlap sums, per thread in k order.
// EXACT: only the bubbles whose stored overlap carries the flag sum again;
// the others take their stored overlap and radius, so the partials are those
// of a pass taking the flagged pairs' values from FP64 from the start.
template <int D, bool GAS, bool FLOW, bool EXACT>
__device__ __forceinline__ void mixed_block(const Consts &c, const Roles &R, const Scratch &S,
                                            size_t cap, int base, int ipt, int n, double h,
                                            double &soa, double &soar, bool &near_any) {
    float step[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) step[a] = fixed_step(c.interval[a], c.wall[a] != 0);
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        float ri = 1.f, ovl = 0.f;
        bool redo = true;
        if (EXACT) {
            ovl = field<float>(S.ov)[i];
            redo = (__float_as_uint(ovl) >> 31) != 0;
        }
        if (redo) {
            bool near = false;
            ovl = mixed_bubble<D, GAS, FLOW, EXACT>(c, R, S, cap, h, step, i, ri, near);
            near_any = near_any || near;
        } else {
            ri = load_rec(gmem(S.q4) + i).r;
        }
        soa += (double)ovl;
        soar += (double)ovl * (1.0 / (double)ri);
    }
}

template <bool FLOW, bool EXACT>
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
        if (gas)
            mixed_block<3, true, FLOW, EXACT>(c, R, S, cap, base, ipt, n, h, soa, soar, near_any);
        else
            mixed_block<3, false, FLOW, EXACT>(c, R, S, cap, base, ipt, n, h, soa, soar, near_any);
    } else {
        if (gas)
            mixed_block<2, true, FLOW, EXACT>(c, R, S, cap, base, ipt, n, h, soa, soar, near_any);
        else
            mixed_block<2, false, FLOW, EXACT>(c, R, S, cap, base, ipt, n, h, soa, soar, near_any);
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
    const int mark = block_sum_i(wave_bit);
    soa = block_sum(soa);
    soar = block_sum(soar);
    if (threadIdx.x == 0) {
        F->s.part[cx.lb * kPartD + 2] = soa;
        F->s.part[cx.lb * kPartD + 3] = soar;
        F->s.part_i[cx.lb * kPartI + 2] = mark;
    }
}

// The flagged bubbles of every foam, in parallel: summed again with the
// values and decisions of the pairs that need them taken from the FP64
// predictions (the same bubbles per thread, in the same order, so every
// stored value is that of a pass taking them so from the start), and their
// blocks' overlap partials formed again.  An unmarked workgroup returns at
// once; in a marked one only the lanes of flagged bubbles walk their lists.
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
    soa = block_sum(soa);
    soar = block_sum(soar);
    if (threadIdx.x == 0) {
        F->s.part[cx.lb * kPartD + 2] = soa;
        F->s.part[cx.lb * kPartD + 3] = soar;
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
