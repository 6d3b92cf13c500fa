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
// such pair moves its bubbles' overlap and dR/dt by O(1).  Within a relative
// band of 2^-16 around either threshold the kernel decides the branch from
// the FP64 predictions, exactly as the FP64 path does; the band is ~15x the
// FP32 error and holds about 1e-5 of the pairs.
#include "dev_util.hpp"
#include "launch.hpp"

namespace cubble_amd {

// Relative half-width of the band around a branch threshold that is decided
// in FP64.
constexpr float kBranchBand = 0x1p-16f;

// The FP64 path's two branch decisions for pair (i, j), from the FP64
// predictions (recomputed as k_correct does): bit 0, the pair overlaps
// ((r1 + r2)^2 >= |d|^2); bit 1, a centre lies inside the other bubble
// (|d|^2 < r1^2 or |d|^2 < r2^2), r1 the radius of the smaller index.
template <int D, bool GAS>
__device__ __attribute__((noinline)) int exact_branches(const Consts &c, const Roles &R,
                                                        double h, int i, int j) {
#pragma clang fp contract(off)
    const D4 pi = predicted<D, GAS, float>(R, h, i);
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

// FLOW: also the imposed-flow term (the neighbour-velocity sum in FP32, the
// flow-box test on the bubble's FP64 prediction, recomputed here as the
// corrector does); same neighbour selection as the FP64 kernel.
template <int D, bool GAS, bool FLOW>
__device__ __forceinline__ void pair_range_mixed(const Consts &c, const Roles &R,
                                                 const Scratch &S, size_t cap, int base, int ipt,
                                                 int n, double h, double &soa, double &soar) {
    const Q4 *__restrict__ q4 = S.q4;
    const int *__restrict__ nnA = S.nn;
    const int *__restrict__ nbr = S.nbr;
    const float f0 = (float)c.f0;
    const float pi_f = (float)kPi;
    float step[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) step[a] = fixed_step(c.interval[a], c.wall[a] != 0);
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        const Q4 qi = load_rec(gmem(q4) + i);
        const float ri_f = qi.r;
        const float inv_ri = __builtin_amdgcn_rcpf(ri_f);
        float acc[3] = {0.f, 0.f, 0.f};
        float fv[3] = {0.f, 0.f, 0.f};
        float ovl = 0.f, dr = 0.f;
        const int m_nn = nnA[i];
        for_neighbours(nbr, q4, cap, i, m_nn, [&](int j, const Q4 &qj) {
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
                if (j > i && fabsf(over) <= kBranchBand * (rad * rad))
                    take = (exact_branches<D, GAS>(c, R, h, i, j) & 1) != 0;
                if (take) {
#pragma unroll
                    for (int a = 0; a < D; ++a) fv[a] += field<float>(R.v[a])[j];
                }
            }
            if (over >= 0.f) {
                const float rs = __builtin_amdgcn_rsqf(m);
                const float s = rs - __builtin_amdgcn_rcpf(rad);
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] += dd[a] * (f0 * s);
                if (GAS) {
                    const float r1 = i < j ? ri_f : rj_f;
                    const float r2 = i < j ? rj_f : ri_f;
                    const float r1sq = r1 * r1, r2sq = r2 * r2;
                    float a2;
                    bool inside = m < r1sq || m < r2sq;
                    if (fabsf(m - r1sq) <= kBranchBand * r1sq ||
                        fabsf(m - r2sq) <= kBranchBand * r2sq)
                        inside = (exact_branches<D, GAS>(c, R, h, i, j) & 2) != 0;
                    if (inside) {
                        a2 = r1sq < r2sq ? r1sq : r2sq;
                    } else {
                        const float dl = r1 - r2;
                        a2 = over * (m - dl * dl) * (0.25f * rs * rs);
                        a2 = a2 < 0.f ? -a2 : a2;
                    }
                    const float oa = 3 == D ? a2 * pi_f : 2.0f * __builtin_amdgcn_sqrtf(a2);
                    ovl += oa;
                    dr += oa * (__builtin_amdgcn_rcpf(rj_f) - inv_ri);
                }
            }
        });
        if (FLOW) add_flow<D>(c, predicted<D, GAS, float>(R, h, i), S.deg[i], fv, acc);
#pragma unroll
        for (int a = 0; a < D; ++a) stream_store(field<float>(R.vp[a]) + i, acc[a]);
        stream_store(field<float>(R.vp[3]) + i, dr);
        stream_store(field<float>(S.ov) + i, ovl);
        soa += (double)ovl;
        soar += (double)ovl * (1.0 / (double)ri_f);
    }
}

// One kernel per FLOW value (the flow variant needs more registers).
template <bool FLOW>
__global__ __launch_bounds__(kBlock) void k_pair_mixed(Foam *foams, const int2 *bt,
                                                       uint32_t flags) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (F->d.halt) return;
    const int n = F->d.count, ipt = F->ipt;
    const Consts c = F->c;
    const Roles R = F->r;
    const Scratch S = F->s;
    const size_t cap = (size_t)F->cap;
    const int base = cx.lb * kBlock * ipt;
    double soa = 0.0, soar = 0.0;
    const bool gas = (flags & kGas) != 0;
    const double h = 0.5 * F->d.dt;
    if (3 == c.dim) {
        if (gas) pair_range_mixed<3, true, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar);
        else pair_range_mixed<3, false, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar);
    } else {
        if (gas) pair_range_mixed<2, true, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar);
        else pair_range_mixed<2, false, FLOW>(c, R, S, cap, base, ipt, n, h, soa, soar);
    }
    soa = block_sum(soa);
    soar = block_sum(soar);
    if (threadIdx.x == 0) {
        S.part[cx.lb * kPartD + 2] = soa;
        S.part[cx.lb * kPartD + 3] = soar;
    }
}

void launch_pair_mixed(const Grid &g, uint32_t flags) {
    if (flags & kFlow)
        hipLaunchKernelGGL(k_pair_mixed<true>, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams,
                           g.bt_bub, flags);
    else
        hipLaunchKernelGGL(k_pair_mixed<false>, dim3(g.nb_bub), dim3(kBlock), 0, g.stream,
                           g.foams, g.bt_bub, flags);
}

}  // namespace cubble_amd
