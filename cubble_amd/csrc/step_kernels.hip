// step_kernels.hip -- the three kernels of one step() attempt.
//
//   k_predict : AB2 predictor + block partials of S_A, S_Ar; writes the
//               32-byte FP64 neighbour record, or in mixed precision the
//               16-byte fixed-point one (reference preIntegrate,
//               src/kernels.cu:40-170)
//   k_pair    : pairwise Durian velocity, overlap area and gas flux, gathered
//               per bubble over its full neighbour list; the last block
//               reduces S_A, S_Ar, S_OA, S_OAr
//               (reference pairwiseInteraction, src/kernels.cu:172-337)
//   k_correct : walls, liquid-mediated gas exchange, trapezoid corrector,
//               error / expansion / radius / volume reductions, deletion
//               marks, path; the last block runs the dt controller and, when
//               committing, rotates the role pointers
//               (reference postIntegrate, src/kernels.cu:339-495, :546-593,
//                and step()/integrate(), src/simulator.cu:192-314)
//
// The reference scatters every pair into both bubbles with FP64 atomics in
// unspecified order.  Here each bubble gathers its own neighbours in
// ascending index order, so results are deterministic, independent of the
// launch shape, and need no atomics.  All reductions are fixed-order trees.
#include <type_traits>

#include "dev_util.hpp"
#include "launch.hpp"

namespace cubble_amd {

// ---------------------------------------------------------------- predictor
// FP64: writes the 32-byte record the FP64 pair kernel gathers.  Mixed
// (FIXED, FP32 velocities): writes only the 16-byte fixed-point record.
template <int D, bool GAS, bool FIXED>
__device__ __forceinline__ void predict_range(const Consts &c, const Roles &R, D4 *__restrict__ p4,
                                              Q4 *__restrict__ q4, int base, int ipt, int n,
                                              double h, double &sa, double &sar) {
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        const D4 pr = predicted<D, GAS, typename std::conditional<FIXED, float, double>::type>(
            R, h, i);
        if (GAS) {
            double apr = 2.0 * kPi;
            if (3 == D) apr *= 2.0 * pr.r;
            sa += apr * pr.r;
            sar += apr;
        }
        if (FIXED) {
            const double p[3] = {pr.x, pr.y, pr.z};
            int32_t qc[3] = {0, 0, 0};
#pragma unroll
            for (int q = 0; q < D; ++q) qc[q] = fixed_code(p[q], c.lbb[q], c.interval[q], c.wall[q] != 0);
            typedef int v4i __attribute__((ext_vector_type(4)));
            const v4i w = {qc[0], qc[1], qc[2], __float_as_int((float)pr.r)};
            stream_store(reinterpret_cast<v4i *>(q4 + i), w);
        } else {
            p4[i] = pr;
        }
    }
}

template <int D, bool GAS>
__device__ __forceinline__ void predict_dispatch(const Consts &c, const Roles &R,
                                                 D4 *__restrict__ p4, Q4 *__restrict__ q4,
                                                 int base, int ipt, int n, double h, double &sa,
                                                 double &sar) {
    if (q4) predict_range<D, GAS, true>(c, R, p4, q4, base, ipt, n, h, sa, sar);
    else predict_range<D, GAS, false>(c, R, p4, q4, base, ipt, n, h, sa, sar);
}

__global__ __launch_bounds__(kBlock) void k_predict(Foam *foams, const int2 *bt, uint32_t flags) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (F->d.halt) return;
    const int n = F->d.count, ipt = F->ipt;
    const double h = 0.5 * F->d.dt;
    const Roles R = F->r;
    const int base = cx.lb * kBlock * ipt;
    double sa = 0.0, sar = 0.0;
    const bool gas = (flags & kGas) != 0;
    D4 *p4 = F->s.p4;
    Q4 *q4 = F->s.q4;
    const Consts c = F->c;
    if (3 == c.dim) {
        if (gas) predict_dispatch<3, true>(c, R, p4, q4, base, ipt, n, h, sa, sar);
        else predict_dispatch<3, false>(c, R, p4, q4, base, ipt, n, h, sa, sar);
    } else {
        if (gas) predict_dispatch<2, true>(c, R, p4, q4, base, ipt, n, h, sa, sar);
        else predict_dispatch<2, false>(c, R, p4, q4, base, ipt, n, h, sa, sar);
    }
    double sums[2] = {sa, sar};
    int none = 0;
    block_reduce_t0(sums, 0u, none);
    if (threadIdx.x == 0) {
        F->s.part[cx.lb * kPartD + 0] = sums[0];
        F->s.part[cx.lb * kPartD + 1] = sums[1];
    }
}

// ---------------------------------------------------------------- pair (FP64)
// FLOW: also the imposed-flow term.  The neighbour-velocity sum takes every
// listed neighbour with a smaller index and the overlapping ones with a
// larger index: the reference adds dxdt[idx1] into idx2 for every listed pair
// (idx1 < idx2) but dxdt[idx2] into idx1 only past its overlap test
// (src/kernels.cu:191-208 vs :281-307).
template <int D, bool GAS, bool FLOW>
__device__ __forceinline__ void pair_range(const Consts &c, const Roles &R, const Scratch &S,
                                           size_t cap, int base, int ipt, int n, double &soa,
                                           double &soar) {
    const D4 *__restrict__ p4 = S.p4;
    const int *__restrict__ nnA = S.nn;
    const int *__restrict__ nbr = S.nbr;
    const double f0 = c.f0;
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        const D4 pi = p4[i];
        const double xi = pi.x, yi = pi.y, zi = pi.z, ri = pi.r;
        const double inv_ri = 1.0 / ri;
        double acc[3] = {0.0, 0.0, 0.0};
        double fv[3] = {0.0, 0.0, 0.0};
        double ovl = 0.0, dr = 0.0;
        const int m_nn = nnA[i];
        for_neighbours(nbr, p4, cap, i, m_nn, [&](int j, const D4 &pj) {
            const double rj = pj.r;
            const double r1 = i < j ? ri : rj;
            const double r2 = i < j ? rj : ri;
            const double radii = r1 + r2;
            double d[3];
            wrapped_diff<D>(c, xi, yi, zi, pj.x, pj.y, pj.z, d);
            const double m = sq_len(d);
            if (FLOW && (j < i || radii * radii >= m)) {
#pragma unroll
                for (int a = 0; a < D; ++a) fv[a] += field<double>(R.v[a])[j];
            }
            if (radii * radii >= m) {
                const double rs = rsq_nr(m);
                const double s = rs - rcp_nr(radii);
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] += d[a] * f0 * s;
                if (GAS) {
                    const double r1sq = r1 * r1;
                    const double r2sq = r2 * r2;
                    double oa = 0;
                    if (m < r1sq || m < r2sq) {
                        oa = r1sq < r2sq ? r1sq : r2sq;
                    } else {
                        oa = 0.5 * (r2sq - r1sq + m);
                        oa *= oa;
                        oa *= rs * rs;  // / m
                        oa = r2sq - oa;
                        oa = oa < 0 ? -oa : oa;
                    }
                    if (3 == D) {
                        oa *= kPi;
                    } else {
                        oa = 2.0 * sqrt(oa);
                    }
                    ovl += oa;
                    dr += oa * (rcp_nr(rj) - inv_ri);
                }
            }
        });
        if (FLOW) add_flow<D>(c, pi, S.deg[i], fv, acc);
#pragma unroll
        for (int a = 0; a < D; ++a) field<double>(R.vp[a])[i] = acc[a];
        field<double>(R.vp[3])[i] = dr;
        field<double>(S.ov)[i] = ovl;
        soa += ovl;
        soar += ovl * inv_ri;
    }
}

// Four consecutive slots q0..q0+3 of a foam's block partials, reduced in one
// pass over the records (each thread takes blocks t, t + kBlock, ... in
// order), and the integer slot qi of the int partials summed alongside when
// qi >= 0.  Bit k of `dmax` makes slot q0 + k a maximum (from 0), else a
// sum.  The results are defined in thread 0 only.
__device__ __forceinline__ void reduce_parts4(const double *part, const int32_t *part_i, int nb,
                                              int q0, unsigned dmax, int qi, double (&r)[4],
                                              int &ri) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    int ai = 0;
#pragma unroll 2
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        const double *p = part + b * kPartD + q0;
        const double v[4] = {p[0], p[1], p[2], p[3]};
#pragma unroll
        for (int k = 0; k < 4; ++k) a[k] = (dmax >> k) & 1u ? fmax(a[k], v[k]) : a[k] + v[k];
        if (qi >= 0) ai += part_i[b * kPartI + qi];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = a[k];
    ri = ai;
    block_reduce_t0(r, dmax, ri);
}

// One kernel per FLOW value: the flow variant needs more registers, which
// would otherwise lower the occupancy of the plain one.
template <bool FLOW>
__global__ __launch_bounds__(kBlock) void k_pair(Foam *foams, const int2 *bt, uint32_t flags) {
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
        if (gas) pair_range<3, true, FLOW>(c, R, S, cap, base, ipt, n, soa, soar);
        else pair_range<3, false, FLOW>(c, R, S, cap, base, ipt, n, soa, soar);
    } else {
        if (gas) pair_range<2, true, FLOW>(c, R, S, cap, base, ipt, n, soa, soar);
        else pair_range<2, false, FLOW>(c, R, S, cap, base, ipt, n, soa, soar);
    }
    double sums[2] = {soa, soar};
    int none = 0;
    block_reduce_t0(sums, 0u, none);
    if (threadIdx.x == 0) {
        S.part[cx.lb * kPartD + 2] = sums[0];
        S.part[cx.lb * kPartD + 3] = sums[1];
    }
}

// One block per foam: S_A, S_Ar, S_OA, S_OAr from the predictor and pair
// partials, in a fixed order.  (A separate launch: on MI355X the eight XCD
// L2 caches are not coherent, so an in-kernel last-block reduction would
// need an agent-scope release -- an L2 write-back -- in every block.)
__global__ __launch_bounds__(kBlock) void k_sums(Foam *foams) {
    Foam *F = foams + blockIdx.x;
    if (F->d.halt) return;
    double r[4];
    int none;
    reduce_parts4(F->s.part, F->s.part_i, F->nblk, 0, 0u, -1, r, none);
    if (threadIdx.x == 0) {
        F->d.s_a = r[0];
        F->d.s_ar = r[1];
        F->d.s_oa = r[2];
        F->d.s_oar = r[3];
    }
}

// ---------------------------------------------------------------- corrector
// reference src/kernels.cu:546-593 (addWallVelocity)
template <int D>
__device__ __forceinline__ void add_wall_velocity(const Consts &c, double xp, double yp, double zp,
                                                  double rad, double &vx, double &vy, double &vz) {
    const double drag = 1.0 - c.wall_drag;
    double xDrag = 1.0, velocity = 0.0;
    auto touches = [&](double x, double low, double high) -> bool {
        double d1 = x - low;
        double d2 = x - high;
        d1 = d1 * d1 < d2 * d2 ? d1 : d2;
        if (rad * rad >= d1 * d1) {
            d2 = d1 < 0.0 ? -1.0 : 1.0;
            d1 *= d2;
            velocity = d2 * c.f0 * (1.0 - d1 / rad);
            return true;
        }
        return false;
    };
    double yDrag = 1.0;
    velocity = 0.0;
    if (c.wall[0] && touches(xp, c.lbb[0], c.tfr[0])) {
        vx += velocity;
        vy *= drag;
        if (3 == D) vz *= drag;
        xDrag = drag;
    }
    velocity = 0.0;
    if (c.wall[1] && touches(yp, c.lbb[1], c.tfr[1])) {
        vx *= drag;
        vy += velocity * xDrag;
        if (3 == D) vz *= drag;
        yDrag = drag;
    }
    velocity = 0.0;
    if (3 == D && c.wall[2] && touches(zp, c.lbb[2], c.tfr[2])) {
        vx *= drag;
        vy *= drag;
        vz += velocity * xDrag * yDrag;
    }
}

struct CorrAcc {
    double s_v, max_err, max_r, max_exp;
    int n_del;
};

// MIXED: no FP64 predicted record is stored (recompute it) and velocities
// are FP32.  WALLS false: no foam of the launch has a wall (the wall code,
// and its registers, compiled out).
template <int D, bool GAS, bool PATH, bool MIXED, bool WALLS>
__device__ __forceinline__ void correct_range(const Consts &c, const Roles &R, const Scratch &S,
                                              const Dyn &dyn, int base, int ipt, int n,
                                              CorrAcc &a) {
    typedef typename std::conditional<MIXED, float, double>::type V;
    const double h = 0.5 * dyn.dt;
    const bool walls = WALLS && (c.wall[0] || c.wall[1] || (3 == D && c.wall[2]));
    const double inv_rho = (dyn.s_ar - dyn.s_oar) / (dyn.s_a - dyn.s_oa);
    const double flux = c.kappa * c.avg_area_in * n / dyn.s_a;  // same for every bubble
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        const D4 pp = MIXED ? predicted<D, GAS, V>(R, h, i) : S.p4[i];
        const double pv[3] = {pp.x, pp.y, pp.z};
        double vp[3];
#pragma unroll
        for (int q = 0; q < D; ++q) vp[q] = (double)field<V>(R.vp[q])[i];
        if (walls) {
            double vz = 3 == D ? vp[2] : 0.0;
            add_wall_velocity<D>(c, pp.x, pp.y, pp.z, pp.r, vp[0], vp[1], vz);
            if (3 == D) vp[2] = vz;
#pragma unroll
            for (int q = 0; q < D; ++q) {
                vp[q] = (double)(V)vp[q];  // the stored value drives the corrector
                stream_store(field<V>(R.vp[q]) + i, (V)vp[q]);
            }
        }
        double e = 0.0, dist = 0.0;
        double xo[3], xc[3];
#pragma unroll
        for (int q = 0; q < D; ++q) {
            const double pred = pv[q];
            xo[q] = R.x[q][i];
            const double corr = xo[q] + h * ((double)field<V>(R.v[q])[i] + vp[q]);
            xc[q] = corr;
            stream_store(R.xp[q] + i, corr);
            e = fmax(e, fabs(pred - corr));
            const double t = corr - (double)field<V>(R.saved[q])[i];
            dist += t * t;
        }
        dist = sqrt(dist);
        int dele = 0;
        if (GAS) {
            double r = pp.r;
            double area = 2.0 * kPi * r;
            if (3 == D) area *= 2.0 * r;
            double vr = inv_rho - rcp_nr(r);
            vr *= flux * (area - (double)field<V>(S.ov)[i]);
            vr += (double)field<V>(R.vp[3])[i];
            vr *= c.k_param * rcp_nr(area);
            stream_store(field<V>(R.vp[3]) + i, (V)vr);
            vr = (double)(V)vr;  // the stored value drives the corrector
            const double pred = r;
            const double corr = R.x[3][i] + h * ((double)field<V>(R.v[3])[i] + vr);
            stream_store(R.xp[3] + i, corr);
            e = fmax(e, fabs(pred - corr));
            dist += corr - (double)field<V>(R.saved[3])[i];
            r = corr;
            if (r > c.min_rad) {
                double vol = r * r;
                if (3 == D) vol *= r;
                a.s_v += vol;
            } else {
                dele = 1;
                a.n_del += 1;
            }
            a.max_r = fmax(a.max_r, r);
        } else {
            stream_store(R.xp[3] + i, pp.r);  // radius unchanged without gas exchange
        }
        stream_store(S.del + i, (uint8_t)dele);
        a.max_exp = fmax(a.max_exp, dist);
        if (PATH) {
            double t = xo[0] - xc[0];
            double pd = t * t;
            t = xo[1] - xc[1];
            pd += t * t;
            if (3 == D) {
                t = xo[2] - xc[2];
                pd += t * t;
            }
            stream_store(R.path_new + i, R.path[i] + sqrt(pd));
        }
        stream_store(field<V>(R.err) + i, (V)fmax(e, (double)field<V>(R.err)[i]));
        a.max_err = fmax(a.max_err, e);
    }
}

// dt controller and commit: reference src/simulator.cu:239-246, :258-321.
__device__ void control(Foam *F, uint32_t flags, const CorrAcc &a) {
    Dyn &d = F->d;
    const Consts &c = F->c;
    d.s_v = a.s_v;
    d.max_err = a.max_err;
    d.max_r = a.max_r;
    d.max_exp = a.max_exp;
    d.n_del = a.n_del;
    d.attempts += 1;
    const double dt = d.dt;
    const bool too_large = a.max_err > c.tol;
    const bool increase = a.max_err < 0.45 * c.tol && dt < 10;
    double ndt = dt;
    if (too_large) {
        ndt = dt * 0.37;
    } else if (increase) {
        ndt = dt * 1.269;
    }
    d.dt = ndt;
    d.accepted = too_large ? 0 : 1;
    if (too_large || !(flags & kCommit)) return;
    const bool gas = (flags & kGas) != 0;
    Roles &R = F->r;
    for (int k = 0; k < 4; ++k) {
        if (2 == k && 2 == c.dim) continue;
        if (3 == k && !gas) continue;
        void *t = R.vo[k];
        R.vo[k] = R.v[k];
        R.v[k] = R.vp[k];
        R.vp[k] = t;
        double *u = R.x[k];
        R.x[k] = R.xp[k];
        R.xp[k] = u;
    }
    if (flags & kPath) {
        double *t = R.path;
        R.path = R.path_new;
        R.path_new = t;
    }
    d.steps += 1;
    d.bubble_steps += (double)d.count;
    d.time_frac += ndt;
    const uint64_t whole = (uint64_t)d.time_frac;
    d.time_int += whole;
    d.time_frac = d.time_frac - (double)whole;
    if (gas) {
        d.max_r_cell = a.max_r;
        d.need_remove = a.n_del > 0 ? 1 : 0;
    }
    d.need_search = a.max_exp >= 0.5 * c.skin ? 1 : 0;
    d.ts_min = fmin(d.ts_min, ndt);
    d.ts_max = fmax(d.ts_max, ndt);
    d.ts_sum += ndt;
    if (stop_reached(d)) d.halt = 1;
}

template <int D, bool M, bool W>
__device__ __forceinline__ void correct_dispatch(const Consts &c, const Roles &R, const Scratch &S,
                                                 const Dyn &dyn, int base, int ipt, int n,
                                                 uint32_t flags, CorrAcc &a) {
    const bool gas = (flags & kGas) != 0, path = (flags & kPath) != 0;
    if (gas) {
        if (path) correct_range<D, true, true, M, W>(c, R, S, dyn, base, ipt, n, a);
        else correct_range<D, true, false, M, W>(c, R, S, dyn, base, ipt, n, a);
    } else {
        if (path) correct_range<D, false, true, M, W>(c, R, S, dyn, base, ipt, n, a);
        else correct_range<D, false, false, M, W>(c, R, S, dyn, base, ipt, n, a);
    }
}

// One kernel per precision (MIXED: FP32 velocity buffers, recomputed
// prediction) and per wall presence (WALLS: some foam of the launch has a
// wall), so that none pays the registers of another.
template <bool MIXED, bool WALLS>
__global__ __launch_bounds__(kBlock) void k_correct(Foam *foams, const int2 *bt, uint32_t flags) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (F->d.halt) return;
    const int n = F->d.count, ipt = F->ipt;
    const Consts c = F->c;
    const Roles R = F->r;
    const Scratch S = F->s;
    const Dyn dyn = F->d;
    const int base = cx.lb * kBlock * ipt;
    CorrAcc a{0.0, 0.0, 0.0, 0.0, 0};
    if (3 == c.dim) correct_dispatch<3, MIXED, WALLS>(c, R, S, dyn, base, ipt, n, flags, a);
    else correct_dispatch<2, MIXED, WALLS>(c, R, S, dyn, base, ipt, n, flags, a);
    // the volume sum, then the maxima of error, radius and expansion
    double r4[4] = {a.s_v, a.max_err, a.max_r, a.max_exp};
    int nd = a.n_del;
    block_reduce_t0(r4, 0xeu, nd);
    if (threadIdx.x == 0) {
        S.part[cx.lb * kPartD + 4] = r4[0];
        S.part[cx.lb * kPartD + 5] = r4[1];
        S.part[cx.lb * kPartD + 6] = r4[2];
        S.part[cx.lb * kPartD + 7] = r4[3];
        S.part_i[cx.lb * kPartI + 0] = nd;
    }
}

// One block per foam: reduce the corrector partials, then the dt controller
// and (when committing) the role rotation.
__global__ __launch_bounds__(kBlock) void k_control(Foam *foams, uint32_t flags) {
    Foam *F = foams + blockIdx.x;
    if (F->d.halt) return;
    // the volume sum, the maxima of error, radius and expansion, the
    // deletion count
    double r[4];
    int nd;
    reduce_parts4(F->s.part, F->s.part_i, F->nblk, 4, 0xeu, 0, r, nd);
    if (threadIdx.x == 0) {
        CorrAcc t;
        t.s_v = r[0];
        t.max_err = r[1];
        t.max_r = r[2];
        t.max_exp = r[3];
        t.n_del = nd;
        control(F, flags, t);
    }
}

// ---------------------------------------------------------------- launches
void launch_predict(const Grid &g, uint32_t flags) {
    hipLaunchKernelGGL(k_predict, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams, g.bt_bub,
                       flags);
}

void launch_pair(const Grid &g, uint32_t flags) {
    if (flags & kFlow)
        hipLaunchKernelGGL(k_pair<true>, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams,
                           g.bt_bub, flags);
    else
        hipLaunchKernelGGL(k_pair<false>, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams,
                           g.bt_bub, flags);
}

void launch_sums(const Grid &g) {
    hipLaunchKernelGGL(k_sums, dim3(g.nfoams), dim3(kBlock), 0, g.stream, g.foams);
}

// AB2 bootstrap of init (reference src/simulator.cu:739-787): the Euler step
// of the predicted record by the pair velocities in vp (kernel euler,
// src/kernels.cu:958-968), then those velocities become the old ones.  The
// record holds the current positions (zero velocities at creation); in mixed
// precision the stepped FP64 position is recoded into the fixed-point record.
__global__ __launch_bounds__(kBlock) void k_euler(Foam *foams, const int2 *bt) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    const int n = F->d.count, ipt = F->ipt, D = F->c.dim;
    const Consts c = F->c;
    const Roles R = F->r;
    const Scratch S = F->s;
    const double ts = F->d.dt;
    const int base = cx.lb * kBlock * ipt;
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        if (S.q4) {
            Q4 q = S.q4[i];
            auto code = [&](int a) {
                const double x = R.x[a][i] + (double)field<float>(R.vp[a])[i] * ts;
                return fixed_code(x, c.lbb[a], c.interval[a], c.wall[a] != 0);
            };
            q.x = code(0);
            q.y = code(1);
            if (3 == D) q.z = code(2);
            S.q4[i] = q;
        } else {
            D4 p = S.p4[i];
            p.x += field<double>(R.vp[0])[i] * ts;
            p.y += field<double>(R.vp[1])[i] * ts;
            if (3 == D) p.z += field<double>(R.vp[2])[i] * ts;
            S.p4[i] = p;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_old_from_pair(Foam *foams, const int2 *bt) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    const int n = F->d.count, ipt = F->ipt, D = F->c.dim;
    const Roles R = F->r;
    const bool mixed = F->s.q4 != nullptr;
    const int base = cx.lb * kBlock * ipt;
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        for (int a = 0; a < D; ++a) {
            if (mixed) field<float>(R.vo[a])[i] = field<float>(R.vp[a])[i];
            else field<double>(R.vo[a])[i] = field<double>(R.vp[a])[i];
        }
    }
}

void launch_bootstrap(const Grid &g, bool mixed) {
    launch_predict(g, 0);  // zero velocities: the record holds the positions
    if (mixed) launch_pair_mixed(g, 0);
    else launch_pair(g, 0);
    hipLaunchKernelGGL(k_euler, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams, g.bt_bub);
    if (mixed) launch_pair_mixed(g, 0);
    else launch_pair(g, 0);
    hipLaunchKernelGGL(k_old_from_pair, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams,
                       g.bt_bub);
}

void launch_correct(const Grid &g, uint32_t flags, bool mixed, bool control) {
    const dim3 grid(g.nb_bub), block(kBlock);
    if (mixed) {
        if (g.walls)
            hipLaunchKernelGGL((k_correct<true, true>), grid, block, 0, g.stream, g.foams,
                               g.bt_bub, flags);
        else
            hipLaunchKernelGGL((k_correct<true, false>), grid, block, 0, g.stream, g.foams,
                               g.bt_bub, flags);
    } else {
        if (g.walls)
            hipLaunchKernelGGL((k_correct<false, true>), grid, block, 0, g.stream, g.foams,
                               g.bt_bub, flags);
        else
            hipLaunchKernelGGL((k_correct<false, false>), grid, block, 0, g.stream, g.foams,
                               g.bt_bub, flags);
    }
    if (control)
        hipLaunchKernelGGL(k_control, dim3(g.nfoams), dim3(kBlock), 0, g.stream, g.foams, flags);
}

}  // namespace cubble_amd
