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
// (:281-307).  FP32 distances (relative error ~1e-7) would decide those two
// branch points differently from FP64 for pairs that close to them, and one
// such pair moves its bubbles' overlap and dR/dt by O(1).  A wave whose
// bubbles met a pair within a relative band of 2^-16 around either threshold
// (the band is ~15x the FP32 error and holds about 1e-5 of the pairs) marks
// itself, and a second kernel sums the marked waves' bubbles again with those
// branches decided from the FP64 predictions, exactly as the FP64 path
// decides them, in parallel over the marked waves of every foam.
#include "dev_util.hpp"
#include "launch.hpp"

namespace cubble_amd {

// Relative half-width of the band around a branch threshold that is decided
// in FP64.
constexpr float kBranchBand = 0x1p-16f;

// The FP64 path's two branch decisions for pair (i, j), from the FP64
// predictions (recomputed as k_correct does; pi is bubble i's): bit 0, the
// pair overlaps ((r1 + r2)^2 >= |d|^2); bit 1, a centre lies inside the
// other bubble (|d|^2 < r1^2 or |d|^2 < r2^2), r1 the radius of the smaller
// index.
template <int D, bool GAS>
__device__ __forceinline__ int exact_branches(const Consts &c, const Roles &R, double h,
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
    return (radii * radii >= m ? 1 : 0) | ((m < r1sq || m < r2sq) ? 2 : 0);
}

struct MixedOut {
    float acc[3], fv[3], ovl, dr;
};

// One bubble's pair sums over its list, ascending.  EXACT = false decides
// every branch in FP32 and reports whether a pair came within the band of a
// threshold; EXACT = true (run only for such bubbles) decides those pairs'
// branches in FP64 instead, with everything else the same.
template <int D, bool GAS, bool FLOW, bool EXACT>
__device__ __forceinline__ bool pair_bubble_mixed(const Consts &c, const Roles &R,
                                                  const Q4 *__restrict__ q4,
                                                  const int *__restrict__ nbr, size_t cap,
                                                  double h, const float step[3], int i,
                                                  const Q4 &qi, int m_nn, MixedOut &o) {
    const float f0 = (float)c.f0;
    const float pi_f = (float)kPi;
    const float ri_f = qi.r;
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
        const float m = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
        const float rad = ri_f + rj_f;
        const float over = rad * rad - m;
        if (FLOW) {
            bool take = j < i || over >= 0.f;
            if (j > i && fabsf(over) <= kBranchBand * (rad * rad)) {
                near = true;
                if (EXACT) take = (exact_branches<D, GAS>(c, R, h, pi, i, j) & 1) != 0;
            }
            if (take) {
#pragma unroll
                for (int a = 0; a < D; ++a) o.fv[a] += field<float>(R.v[a])[j];
            }
        }
        if (over >= 0.f) {
            const float rs = __builtin_amdgcn_rsqf(m);
            const float s = rs - __builtin_amdgcn_rcpf(rad);
#pragma unroll
            for (int a = 0; a < D; ++a) o.acc[a] += dd[a] * (f0 * s);
            if (GAS) {
                // a centre inside the other bubble: |d|^2 < r1^2 or r2^2, that
                // is below the larger square; the product form is symmetric in
                // the radii, so their order does not matter
                const float sq_i = ri_f * ri_f, sq_j = rj_f * rj_f;
                const float big = fmaxf(sq_i, sq_j);
                float a2;
                bool inside = m < big;
                if (fabsf(m - big) <= kBranchBand * big) {
                    near = true;
                    if (EXACT) inside = (exact_branches<D, GAS>(c, R, h, pi, i, j) & 2) != 0;
                }
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
// within the band of a branch threshold.  FLOW: also the imposed-flow term
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
    stream_store(field<float>(S.ov) + i, o.ovl);
    ri = qi.r;
    return o.ovl;
}

// The bubbles of bubble-space block lb (thread t takes base + k kBlock + t,
// k < ipt): pair terms and the block's overlap sums, per thread in k order.
// EXACT: only the threads with `redo` sum again; the others take their
// bubbles' stored overlap and radius, so the partials are those of a pass
// deciding the marked waves' branches in FP64 from the start.
template <int D, bool GAS, bool FLOW, bool EXACT>
__device__ __forceinline__ void mixed_block(const Consts &c, const Roles &R, const Scratch &S,
                                            size_t cap, int base, int ipt, int n, double h,
                                            bool redo, double &soa, double &soar,
                                            bool &near_any) {
    float step[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) step[a] = fixed_step(c.interval[a], c.wall[a] != 0);
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        float ri = 1.f, ovl;
        if (!EXACT || redo) {
            bool near = false;
            ovl = mixed_bubble<D, GAS, FLOW, EXACT>(c, R, S, cap, h, step, i, ri, near);
            near_any = near_any || near;
        } else {
            ovl = field<float>(S.ov)[i];
            ri = load_rec(gmem(S.q4) + i).r;
        }
        soa += (double)ovl;
        soar += (double)ovl * (1.0 / (double)ri);
    }
}

template <bool FLOW, bool EXACT>
__device__ __forceinline__ void mixed_block_dispatch(const Foam *F, int lb, uint32_t flags,
                                                     bool redo, double &soa, double &soar,
                                                     bool &near_any) {
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
            mixed_block<3, true, FLOW, EXACT>(c, R, S, cap, base, ipt, n, h, redo, soa, soar,
                                              near_any);
        else
            mixed_block<3, false, FLOW, EXACT>(c, R, S, cap, base, ipt, n, h, redo, soa, soar,
                                               near_any);
    } else {
        if (gas)
            mixed_block<2, true, FLOW, EXACT>(c, R, S, cap, base, ipt, n, h, redo, soa, soar,
                                              near_any);
        else
            mixed_block<2, false, FLOW, EXACT>(c, R, S, cap, base, ipt, n, h, redo, soa, soar,
                                               near_any);
    }
}

// One kernel per FLOW value (the flow variant needs more registers).  Every
// branch is decided in FP32; each wave that met a pair within the band of a
// threshold sets its bit in the block's mark (block partial slot 2).
template <bool FLOW>
__global__ __launch_bounds__(kBlock) void k_pair_mixed(Foam *foams, const int2 *bt,
                                                       uint32_t flags) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (F->d.halt) return;
    double soa = 0.0, soar = 0.0;
    bool near = false;
    mixed_block_dispatch<FLOW, false>(F, cx.lb, flags, true, soa, soar, near);
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

// The marked waves of every foam, in parallel: their bubbles summed again
// with the threshold branches decided in FP64 (the same bubbles per thread,
// in the same order, so every stored value is that of a pass deciding them
// so from the start), and the block's overlap partials formed again.  An
// unmarked workgroup returns at once.
template <bool FLOW>
__global__ __launch_bounds__(kBlock) void k_fix_mixed(Foam *foams, const int2 *bt,
                                                      uint32_t flags) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (F->d.halt) return;
    const int mark = F->s.part_i[cx.lb * kPartI + 2];
    if (!mark) return;
    const bool redo = ((mark >> (threadIdx.x >> 6)) & 1) != 0;
    double soa = 0.0, soar = 0.0;
    bool near = false;
    mixed_block_dispatch<FLOW, true>(F, cx.lb, flags, redo, soa, soar, near);
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
