// pair_mixed.hip -- mixed-precision pairwise kernel.
//
// Same gather structure as the FP64 pair kernel (ascending neighbour order,
// deterministic), but only the relative coordinate d = wrap(x_i - x_j) is
// formed in FP64; the Durian velocity, overlap area and gas flux are then
// evaluated in FP32 with hardware reciprocal square root / reciprocal, and
// accumulated per bubble in FP32.  Positions, radii and everything the
// corrector does stay FP64.  The overlap area uses the product form
//     a = (R^2 - |d|^2) (|d|^2 - (r1 - r2)^2) / (4 |d|^2),
// algebraically equal to the reference's r2^2 - ((r2^2 - r1^2 + |d|^2)/2)^2/|d|^2
// (src/kernels.cu:240-249) but without its cancellation near tangency; the
// factor R^2 - |d|^2 is formed in FP64 before rounding to FP32.
#include "dev_util.hpp"
#include "launch.hpp"

namespace cubble_amd {

template <int D, bool GAS>
__device__ __forceinline__ void pair_range_mixed(const Consts &c, const Roles &R,
                                                 const Scratch &S, size_t cap, int base, int ipt,
                                                 int n, double &soa, double &soar) {
    const D4 *__restrict__ p4 = S.p4;
    const int *__restrict__ nnA = S.nn;
    const int *__restrict__ nbr = S.nbr;
    const float f0 = (float)c.f0;
    const float pi_f = (float)kPi;
    double half[3], L[3];
    bool per[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        L[a] = c.interval[a];
        half[a] = 0.5 * c.interval[a];
        per[a] = !c.wall[a];
    }
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        const D4 pi = p4[i];
        const double xi = pi.x, yi = pi.y, zi = pi.z, ri = pi.r;
        const float ri_f = (float)ri;
        const float inv_ri = __frcp_rn(ri_f);
        float acc[3] = {0.f, 0.f, 0.f};
        float ovl = 0.f, dr = 0.f;
        const int m_nn = nnA[i];
        for_neighbours(nbr, p4, cap, i, m_nn, [&](int j, const D4 &pj) {
            const double rj = pj.r;
            double dd[3] = {xi - pj.x, yi - pj.y, D == 3 ? zi - pj.z : 0.0};
#pragma unroll
            for (int a = 0; a < D; ++a) {
                if (per[a] && fabs(dd[a]) > half[a]) dd[a] -= dd[a] < 0 ? -L[a] : L[a];
            }
            const double m64 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
            const double rad64 = ri + rj;
            const double over64 = rad64 * rad64 - m64;
            if (over64 >= 0.0) {
                const float m = (float)m64;
                const float rs = __frsqrt_rn(m);
                const float s = rs - __frcp_rn((float)rad64);
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] += (float)dd[a] * (f0 * s);
                if (GAS) {
                    const float rj_f = (float)rj;
                    const float r1 = i < j ? ri_f : rj_f;
                    const float r2 = i < j ? rj_f : ri_f;
                    const float r1sq = r1 * r1, r2sq = r2 * r2;
                    float a2;
                    if (m < r1sq || m < r2sq) {
                        a2 = r1sq < r2sq ? r1sq : r2sq;
                    } else {
                        const float dl = r1 - r2;
                        a2 = (float)over64 * (m - dl * dl) * (0.25f * rs * rs);
                        a2 = a2 < 0.f ? -a2 : a2;
                    }
                    const float oa = 3 == D ? a2 * pi_f : 2.0f * __fsqrt_rn(a2);
                    ovl += oa;
                    dr += oa * (__frcp_rn(rj_f) - inv_ri);
                }
            }
        });
#pragma unroll
        for (int a = 0; a < D; ++a) R.vp[a][i] = (double)acc[a];
        R.vp[3][i] = (double)dr;
        S.ov[i] = (double)ovl;
        soa += (double)ovl;
        soar += (double)ovl * (1.0 / ri);
    }
}

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
    if (3 == c.dim) {
        if (gas) pair_range_mixed<3, true>(c, R, S, cap, base, ipt, n, soa, soar);
        else pair_range_mixed<3, false>(c, R, S, cap, base, ipt, n, soa, soar);
    } else {
        if (gas) pair_range_mixed<2, true>(c, R, S, cap, base, ipt, n, soa, soar);
        else pair_range_mixed<2, false>(c, R, S, cap, base, ipt, n, soa, soar);
    }
    soa = block_sum(soa);
    soar = block_sum(soar);
    if (threadIdx.x == 0) {
        S.part[cx.lb * kPartD + 2] = soa;
        S.part[cx.lb * kPartD + 3] = soar;
    }
}

void launch_pair_mixed(const Grid &g, uint32_t flags) {
    hipLaunchKernelGGL(k_pair_mixed, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams, g.bt_bub,
                       flags);
}

}  // namespace cubble_amd
