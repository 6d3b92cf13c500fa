/*
 * cubble_oracle.cpp -- CPU FP64 restatement of cubble's per-timestep path.
 *
 * TEST INFRASTRUCTURE ONLY (see cubble_oracle.h).  Every function cites the
 * reference file:line it restates; paths are relative to the KJLankinen/cubble
 * source tree.  Build with -ffp-contract=off (no fused multiply-add) so that
 * every floating-point operation rounds exactly as written.
 *
 * Deliberate, documented deviations from the reference (all of which make the
 * reference's race-dependent behaviour deterministic):
 *   - pair contributions are summed per bubble over its full neighbour list in
 *     ascending neighbour index (the reference adds them with atomics in
 *     unspecified order, src/kernels.cu:172-337);
 *   - rsqrt(m) is computed as 1.0 / sqrt(m) (both correctly rounded);
 *   - the cell counting sort is stable (src/kernels.cu:609-615 is not);
 *   - the neighbour list holds both directions; the reference's half list
 *     (src/kernels.cu:617-701) is the subset with i < j;
 *   - compaction fills holes from the tail in ascending hole / descending
 *     survivor order over a sorted deletion list (the reference's list order
 *     comes from an atomic append, src/kernels.cu:425, :775-936);
 *   - the per-bubble error keeps that bubble's own maximum (quirk of
 *     src/kernels.cu:455 keeps the thread's running maximum);
 *   - rcbrt(S_V) is 1.0 / ora_cbrt(S_V), a Newton cube root built only from
 *     correctly rounded operations, so that CPU and GPU agree bit for bit.
 */
#include "cubble_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* reference src/util.cuh:63 */
const double kPi = 3.1415926535897932384626433832795028841971693993;

int g_threads = 1;

struct Foam {
    ora_consts c{};
    int32_t n = 0;
    std::vector<double> x[4], v[4], vo[4], saved[4], path, err;
    std::vector<int32_t> wc[3], idx;
    std::vector<int32_t> nbr_off, nbr;
    /* list length at the last search: the reference's numNeighbors
     * (src/kernels.cu:713-719), which a removal moves with the bubble
     * (src/kernels.cu:820-827) but does not recount */
    std::vector<int32_t> deg;
    /* outputs of the last attempt */
    std::vector<double> xp[4], vp[4], ov, path_new;
    std::vector<uint8_t> del;
    ora_scalars s{};
};

/* reference src/kernels.cu:48-50 */
inline double predict(double p, double v, double vo, double ts) {
    return p + 0.5 * ts * (3.0 * v - vo);
}

/* reference src/kernels.cu:1094-1114 (wrappedDifference) */
inline void wrapped_diff(const ora_consts &c, double x1, double y1, double z1,
                         double x2, double y2, double z2, double d[3]) {
    double d1[3] = {x1 - x2, y1 - y2, 0.0};
    if (3 == c.dim) d1[2] = z1 - z2;
    d[0] = d1[0];
    d[1] = d1[1];
    d[2] = d1[2];
    for (int k = 0; k < c.dim; ++k) {
        const double a = d1[k] < 0 ? -d1[k] : d1[k];
        const double t = c.interval[k] - a;
        if (!c.wall[k] && t * t < d1[k] * d1[k]) d[k] = t * (d1[k] < 0 ? 1.0 : -1.0);
    }
}

/* reference src/vec.h:41-49 (getSquaredLength, summed x, y, z in order) */
inline double sq_len(const double d[3]) {
    double t = 0;
    t += d[0] * d[0];
    t += d[1] * d[1];
    t += d[2] * d[2];
    return t;
}

/* reference src/kernels.cu:497-544 (addFlowVelocity), in its own terms:
 * multiplier = 1 / numNeighbors, the inside test per axis, and the term
 * !inside * multiplier * flowV + flowVel * inside added to the predicted
 * velocity (before the walls). */
void add_flow_velocity(const ora_consts &c, double xi, double yi, double zi, double rp,
                       int32_t num_neighbors, const double flow_v[3], double dxdtp[3]) {
    double multiplier = num_neighbors;
    multiplier = (multiplier > 0 ? 1.0 / multiplier : 0.0);
    const double ri_sq = rp * rp;
    const double x[3] = {xi, yi, zi};
    int32_t inside = 1;
    for (int k = 0; k < c.dim; ++k) {
        inside *= (int32_t)((x[k] < c.flow_tfr[k] && x[k] > c.flow_lbb[k]) ||
                            ((c.flow_lbb[k] - x[k]) * (c.flow_lbb[k] - x[k]) <= ri_sq) ||
                            ((c.flow_tfr[k] - x[k]) * (c.flow_tfr[k] - x[k]) <= ri_sq));
    }
    for (int k = 0; k < c.dim; ++k)
        dxdtp[k] += !inside * multiplier * flow_v[k] + c.flow_vel[k] * inside;
}

/* reference src/kernels.cu:546-593 (addWallVelocity) */
void add_wall_velocity(const ora_consts &c, double xp, double yp, double zp, double rad,
                       double &vx, double &vy, double &vz) {
    const double drag = 1.0 - c.wall_drag;
    double xDrag = 1.0, yDrag = 1.0, velocity = 0.0;
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
    velocity = 0.0;
    if (c.wall[0] && touches(xp, c.lbb[0], c.tfr[0])) {
        vx += velocity;
        vy *= drag;
        if (3 == c.dim) vz *= drag;
        xDrag = drag;
    }
    velocity = 0.0;
    if (c.wall[1] && touches(yp, c.lbb[1], c.tfr[1])) {
        vx *= drag;
        vy += velocity * xDrag;
        if (3 == c.dim) vz *= drag;
        yDrag = drag;
    }
    velocity = 0.0;
    if (3 == c.dim && c.wall[2] && touches(zp, c.lbb[2], c.tfr[2])) {
        vx *= drag;
        vy *= drag;
        vz += velocity * xDrag * yDrag;
    }
}

/* Newton cube root from correctly rounded operations only; stands in for
 * rcbrt in the volume fix (reference src/kernels.cu:861-867). */
double det_cbrt(double a) {
    if (!(a > 0.0) || std::isinf(a)) return a;
    int e = 0;
    double m = std::frexp(a, &e);           /* a = m * 2^e, m in [0.5, 1) */
    int r = e % 3;
    if (r < 0) r += 3;
    m = std::ldexp(m, r);                   /* m in [0.5, 4) */
    e -= r;
    double y = 1.0;
    for (int it = 0; it < 8; ++it) {
        const double y2 = y * y;
        const double f = y2 * y - m;
        const double g = 3.0 * y2;
        y = y - f / g;
    }
    return std::ldexp(y, e / 3);
}

void resize_attempt(Foam &f) {
    const size_t n = f.n;
    for (int k = 0; k < 4; ++k) {
        f.xp[k].assign(n, 0.0);
        f.vp[k].assign(n, 0.0);
    }
    f.ov.assign(n, 0.0);
    f.path_new.assign(n, 0.0);
    f.del.assign(n, 0);
}

/* pairwise velocity and gas flux of the predicted state f.xp into f.vp and
 * f.ov, gathered per bubble: src/kernels.cu:172-337 (pair math :210-280,
 * flow :191-208, :281-307), :1094-1114 (wrap), vec.h:41-49 */
void pair_step(Foam &f, bool gas, bool flow, double &s_oa, double &s_oar) {
    const ora_consts &c = f.c;
    const int32_t n = f.n;
    const int D = c.dim;
#pragma omp parallel for reduction(+ : s_oa, s_oar) if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        double acc[3] = {0.0, 0.0, 0.0};
        double flow_v[3] = {0.0, 0.0, 0.0};
        double ovl = 0.0, dr = 0.0;
        const double ri = f.xp[3][i];
        for (int32_t q = f.nbr_off[i]; q < f.nbr_off[i + 1]; ++q) {
            const int32_t j = f.nbr[q];
            const int32_t a = i < j ? i : j;
            const int32_t b = i < j ? j : i;
            const double r1 = f.xp[3][a];
            const double r2 = f.xp[3][b];
            const double radii = r1 + r2;
            double d[3];
            wrapped_diff(c, f.xp[0][i], f.xp[1][i], 3 == D ? f.xp[2][i] : 0.0, f.xp[0][j],
                         f.xp[1][j], 3 == D ? f.xp[2][j] : 0.0, d);
            const double m = sq_len(d);
            /* flow velocities: the pair (a, b), a < b, adds v[a] into b for
             * every listed pair (src/kernels.cu:191-208) but v[b] into a only
             * past the overlap test (:281-307) */
            if (flow && (i == b || radii * radii >= m))
                for (int k = 0; k < D; ++k) flow_v[k] += f.v[k][j];
            if (radii * radii >= m) {
                const double s = 1.0 / std::sqrt(m) - 1.0 / radii;
                for (int k = 0; k < D; ++k) acc[k] += d[k] * c.f0 * s;
                if (gas) {
                    const double r1sq = r1 * r1;
                    const double r2sq = r2 * r2;
                    double oa = 0;
                    if (m < r1sq || m < r2sq) {
                        oa = r1sq < r2sq ? r1sq : r2sq;
                    } else {
                        oa = 0.5 * (r2sq - r1sq + m);
                        oa *= oa;
                        oa /= m;
                        oa = r2sq - oa;
                        oa = oa < 0 ? -oa : oa;
                    }
                    if (3 == D) {
                        oa *= kPi;
                    } else {
                        oa = 2.0 * std::sqrt(oa);
                    }
                    ovl += oa;
                    dr += oa * (1.0 / f.xp[3][j] - 1.0 / ri);
                }
            }
        }
        if (flow)
            add_flow_velocity(c, f.xp[0][i], f.xp[1][i], 3 == D ? f.xp[2][i] : 0.0, f.xp[3][i],
                              f.deg[i], flow_v, acc);
        for (int k = 0; k < D; ++k) f.vp[k][i] = acc[k];
        f.vp[3][i] = dr;
        f.ov[i] = ovl;
        s_oa += ovl;
        s_oar += ovl * (1.0 / ri);
    }
}

/* One step() attempt: reference src/simulator.cu:192-249 with kernels
 * preIntegrate (src/kernels.cu:40-170), pairwiseInteraction (:172-337) and
 * postIntegrate (:339-495). */
int attempt(Foam &f, bool gas, bool incr_path, bool flow) {
    const ora_consts &c = f.c;
    const int32_t n = f.n;
    const int D = c.dim;
    const double ts = f.s.dt;
    resize_attempt(f);

    /* predictor + S_A, S_Ar: src/kernels.cu:48-113 */
    double s_a = 0.0, s_ar = 0.0;
#pragma omp parallel for reduction(+ : s_a, s_ar) if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        for (int k = 0; k < D; ++k) f.xp[k][i] = predict(f.x[k][i], f.v[k][i], f.vo[k][i], ts);
        if (gas) {
            const double rp = predict(f.x[3][i], f.v[3][i], f.vo[3][i], ts);
            f.xp[3][i] = rp;
            double apr = 2.0 * kPi;
            if (3 == D) apr *= 2.0 * rp;
            s_a += apr * rp;
            s_ar += apr;
        } else {
            f.xp[3][i] = f.x[3][i];
        }
    }

    double s_oa = 0.0, s_oar = 0.0;
    pair_step(f, gas, flow, s_oa, s_oar);

    /* walls, liquid-mediated gas exchange, corrector, error, expansion,
     * deletion marks, volume, path: src/kernels.cu:356-456, :487-495 */
    double s_v = 0.0, max_err = 0.0, max_r = 0.0, max_exp = 0.0;
    int32_t n_del = 0;
    const double inv_rho = (s_ar - s_oar) / (s_a - s_oa);
#pragma omp parallel for reduction(+ : s_v, n_del) reduction(max : max_err, max_r, max_exp) if (g_threads > 1) num_threads(g_threads) schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        if (c.wall[0] || c.wall[1] || (3 == D && c.wall[2])) {
            double vz = 3 == D ? f.vp[2][i] : 0.0;
            add_wall_velocity(c, f.xp[0][i], f.xp[1][i], 3 == D ? f.xp[2][i] : 0.0, f.xp[3][i],
                              f.vp[0][i], f.vp[1][i], vz);
            if (3 == D) f.vp[2][i] = vz;
        }
        double e = 0.0, dist = 0.0;
        for (int k = 0; k < D; ++k) {
            const double pred = f.xp[k][i];
            const double corr = f.x[k][i] + 0.5 * ts * (f.v[k][i] + f.vp[k][i]);
            f.xp[k][i] = corr;
            e = std::fmax(e, std::fabs(pred - corr));
            const double t = corr - f.saved[k][i];
            dist += t * t;
        }
        dist = std::sqrt(dist);
        if (gas) {
            double r = f.xp[3][i];
            double area = 2.0 * kPi * r;
            if (3 == D) area *= 2.0 * r;
            double vr = inv_rho - 1.0 / r;
            vr *= c.kappa * c.avg_area_in * n / s_a * (area - f.ov[i]);
            vr += f.vp[3][i];
            vr *= c.k_param / area;
            f.vp[3][i] = vr;
            const double pred = f.xp[3][i];
            const double corr = f.x[3][i] + 0.5 * ts * (f.v[3][i] + vr);
            f.xp[3][i] = corr;
            e = std::fmax(e, std::fabs(pred - corr));
            dist += corr - f.saved[3][i];
            r = corr;
            if (r > c.min_rad) {
                double vol = r * r;
                if (3 == D) vol *= r;
                s_v += vol;
            } else {
                f.del[i] = 1;
                n_del += 1;
            }
            max_r = std::fmax(max_r, r);
        } else {
            f.vp[3][i] = 0.0;
        }
        max_exp = std::fmax(max_exp, dist);
        if (incr_path) {
            double t = f.x[0][i] - f.xp[0][i];
            double pd = t * t;
            t = f.x[1][i] - f.xp[1][i];
            pd += t * t;
            if (3 == D) {
                t = f.x[2][i] - f.xp[2][i];
                pd += t * t;
            }
            f.path_new[i] = f.path[i] + std::sqrt(pd);
        }
        f.err[i] = std::fmax(e, f.err[i]);
        max_err = std::fmax(max_err, e);
    }

    ora_scalars &s = f.s;
    s.s_a = s_a;
    s.s_ar = s_ar;
    s.s_oa = s_oa;
    s.s_oar = s_oar;
    s.s_v = s_v;
    s.max_err = max_err;
    s.max_r = gas ? max_r : 0.0;
    s.max_exp = max_exp;
    s.n_del = n_del;
    s.attempts += 1;

    /* dt control: src/simulator.cu:239-246 */
    const bool too_large = max_err > c.tol;
    const bool increase = max_err < 0.45 * c.tol && s.dt < 10;
    if (too_large) {
        s.dt *= 0.37;
    } else if (increase) {
        s.dt *= 1.269;
    }
    s.accepted = too_large ? 0 : 1;
    s.need_search = (!too_large && max_exp >= 0.5 * c.skin) ? 1 : 0;
    return s.accepted;
}

/* Commit of an accepted attempt: src/simulator.cu:258-314 */
void commit(Foam &f, bool gas, bool incr_path) {
    const int D = f.c.dim;
    for (int k = 0; k < D; ++k) {
        f.vo[k].swap(f.v[k]);
        f.v[k] = f.vp[k];
        f.x[k] = f.xp[k];
    }
    if (gas) {
        f.vo[3].swap(f.v[3]);
        f.v[3] = f.vp[3];
        f.x[3] = f.xp[3];
    }
    if (incr_path) f.path = f.path_new;
    ora_scalars &s = f.s;
    s.steps += 1;
    s.bubble_steps += (double)f.n;
    s.time_frac += s.dt;
    s.time_int += (uint64_t)s.time_frac;
    s.time_frac = s.time_frac - (uint64_t)s.time_frac;
}

/* Cell grid: src/simulator.cu:69-75, clamped to >= 1 per axis and to the
 * cell capacity by shrinking the largest axis. */
void cell_dims(const Foam &f, int32_t cd[3]) {
    const ora_consts &c = f.c;
    const double edge = 2 * (f.s.max_r_cell + c.skin);
    for (int k = 0; k < 3; ++k) {
        double q = k < c.dim ? std::floor(c.interval[k] / edge) : 1.0;
        if (!(q >= 1.0)) q = 1.0;
        if (q > 1.0e6) q = 1.0e6;
        cd[k] = (int32_t)q;
    }
    if (2 == c.dim) cd[2] = 1;
    while ((int64_t)cd[0] * cd[1] * cd[2] > (int64_t)c.cell_cap) {
        int big = 0;
        for (int k = 1; k < 3; ++k)
            if (cd[k] > cd[big]) big = k;
        cd[big] -= 1;
    }
}

/* Cell index of a position: src/kernels.cu:1191-1208, with each axis index
 * clamped into [0, cellDim) (the reference reads out of bounds when a
 * coordinate rounds onto tfr). */
inline int32_t cell_of(const ora_consts &c, const int32_t cd[3], double x, double y, double z) {
    const double p[3] = {x, y, z};
    int32_t id[3] = {0, 0, 0};
    for (int k = 0; k < c.dim; ++k) {
        double q = std::floor(cd[k] * (p[k] - c.lbb[k]) / c.interval[k]);
        int32_t v = q < 0.0 ? 0 : (q >= (double)cd[k] ? cd[k] - 1 : (int32_t)q);
        id[k] = v;
    }
    return id[2] * cd[0] * cd[1] + id[1] * cd[0] + id[0];
}

template <typename T> void permute(std::vector<T> &a, const std::vector<int32_t> &perm) {
    std::vector<T> b(perm.size());
    for (size_t k = 0; k < perm.size(); ++k) b[k] = a[perm[k]];
    a.swap(b);
}

/* searchNeighbors: src/simulator.cu:60-178 with wrapOverPeriodicBoundaries
 * (src/kernels.cu:992-1011), cellByPosition / scan / indexByCell
 * (:598-615), reorganizeByIndex (:721-770), neighborSearch + comparePair
 * (:617-701), getNeighborCellIndex (:1116-1189). */
void search(Foam &f) {
    const ora_consts &c = f.c;
    const int D = c.dim;
    const int32_t n = f.n;
    ora_scalars &s = f.s;
    s.searches += 1;

    /* wrap positions into the box on periodic axes */
    for (int32_t i = 0; i < n; ++i) {
        for (int k = 0; k < D; ++k) {
            if (c.wall[k]) continue;
            const double xv = f.x[k][i];
            const int32_t mult = xv < c.lbb[k] ? 1 : (xv > c.tfr[k] ? -1 : 0);
            f.x[k][i] = xv + (c.tfr[k] - c.lbb[k]) * (double)mult;
            f.wc[k][i] -= mult;
        }
    }

    int32_t cd[3];
    cell_dims(f, cd);
    s.cell_dim[0] = cd[0];
    s.cell_dim[1] = cd[1];
    s.cell_dim[2] = cd[2];
    const int32_t ncell = cd[0] * cd[1] * cd[2];

    /* stable counting sort of bubbles by cell */
    std::vector<int32_t> cell(n), start(ncell + 1, 0);
    for (int32_t i = 0; i < n; ++i) {
        cell[i] = cell_of(c, cd, f.x[0][i], f.x[1][i], 3 == D ? f.x[2][i] : 0.0);
        start[cell[i] + 1] += 1;
    }
    for (int32_t q = 0; q < ncell; ++q) start[q + 1] += start[q];
    std::vector<int32_t> cursor(start.begin(), start.end() - 1), perm(n);
    for (int32_t i = 0; i < n; ++i) perm[cursor[cell[i]]++] = i;

    /* reorder every persistent per-bubble array; saved <- current */
    for (int k = 0; k < 4; ++k) {
        if (k < D || 3 == k) permute(f.x[k], perm);
        permute(f.v[k], perm);
        permute(f.vo[k], perm);
    }
    permute(f.path, perm);
    permute(f.err, perm);
    for (int k = 0; k < 3; ++k) permute(f.wc[k], perm);
    permute(f.idx, perm);
    for (int k = 0; k < 4; ++k) f.saved[k] = f.x[k];
    std::vector<int32_t> sorted_cell(n);
    for (int32_t i = 0; i < n; ++i) sorted_cell[i] = cell[perm[i]];

    /* full neighbour lists, ascending j: visit the de-duplicated 3^D cell
     * neighbourhood in ascending cell index */
    auto neighbours_of = [&](int32_t i, std::vector<int32_t> &out) {
        out.clear();
        const int32_t ci = sorted_cell[i];
        const int32_t cx = ci % cd[0];
        const int32_t cy = (ci / cd[0]) % cd[1];
        const int32_t cz = ci / (cd[0] * cd[1]);
        int32_t cells[27];
        int nc = 0;
        const int zr = 3 == D ? 1 : 0;
        for (int dz = -zr; dz <= zr; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int32_t q[3] = {cx + dx, cy + dy, cz + dz};
                    bool ok = true;
                    for (int k = 0; k < 3; ++k) {
                        if (k >= D) continue;
                        if (c.wall[k]) {
                            if (q[k] < 0 || q[k] >= cd[k]) ok = false;
                        } else {
                            q[k] = (q[k] + cd[k]) % cd[k];
                        }
                    }
                    if (ok) cells[nc++] = q[2] * cd[0] * cd[1] + q[1] * cd[0] + q[0];
                }
        std::sort(cells, cells + nc);
        nc = (int)(std::unique(cells, cells + nc) - cells);
        const double xi = f.x[0][i], yi = f.x[1][i], zi = 3 == D ? f.x[2][i] : 0.0;
        const double ri = f.x[3][i];
        for (int q = 0; q < nc; ++q) {
            for (int32_t j = start[cells[q]]; j < start[cells[q] + 1]; ++j) {
                if (j == i) continue;
                const double maxd = ri + f.x[3][j] + c.skin;
                double d[3];
                wrapped_diff(c, xi, yi, zi, f.x[0][j], f.x[1][j], 3 == D ? f.x[2][j] : 0.0, d);
                if (sq_len(d) < maxd * maxd) out.push_back(j);
            }
        }
    };

    std::vector<int32_t> counts(n, 0);
#pragma omp parallel if (g_threads > 1) num_threads(g_threads)
    {
        std::vector<int32_t> tmp;
        tmp.reserve(128);
#pragma omp for schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            neighbours_of(i, tmp);
            counts[i] = (int32_t)tmp.size();
        }
    }
    f.nbr_off.assign(n + 1, 0);
    for (int32_t i = 0; i < n; ++i) f.nbr_off[i + 1] = f.nbr_off[i] + counts[i];
    f.nbr.assign(f.nbr_off[n], 0);
#pragma omp parallel if (g_threads > 1) num_threads(g_threads)
    {
        std::vector<int32_t> tmp;
        tmp.reserve(128);
#pragma omp for schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            neighbours_of(i, tmp);
            std::copy(tmp.begin(), tmp.end(), f.nbr.begin() + f.nbr_off[i]);
        }
    }
    s.num_pairs = f.nbr_off[n] / 2;
    f.deg.assign(counts.begin(), counts.end());
}

/* removeBubbles: src/simulator.cu:180-190 with swapDataCountPairs
 * (src/kernels.cu:775-857) and addVolumeFixPairs (:859-936), restated with
 * a sorted deletion list. */
void remove_bubbles(Foam &f) {
    const ora_consts &c = f.c;
    const int32_t n = f.n;
    std::vector<int32_t> dels;
    for (int32_t i = 0; i < n; ++i)
        if (f.del[i]) dels.push_back(i);
    const int32_t nd = (int32_t)dels.size();
    if (0 == nd) return;
    const int32_t nnew = n - nd;
    std::vector<int32_t> holes, tail;
    for (int32_t d : dels)
        if (d < nnew) holes.push_back(d);
    for (int32_t k = n - 1; k >= nnew; --k)
        if (!f.del[k]) tail.push_back(k);
    /* new index of every old index (-1 = deleted) */
    std::vector<int32_t> newidx(n);
    for (int32_t i = 0; i < n; ++i) newidx[i] = f.del[i] ? -1 : i;
    std::vector<std::vector<int32_t>> lists(nnew);
    for (int32_t i = 0; i < nnew; ++i)
        lists[i].assign(f.nbr.begin() + f.nbr_off[i], f.nbr.begin() + f.nbr_off[i + 1]);
    for (size_t h = 0; h < holes.size(); ++h) {
        const int32_t dst = holes[h], src = tail[h];
        for (int k = 0; k < 4; ++k) {
            f.x[k][dst] = f.x[k][src];
            f.v[k][dst] = f.v[k][src];
            f.vo[k][dst] = f.vo[k][src];
            f.saved[k][dst] = f.saved[k][src];
        }
        f.path[dst] = f.path[src];
        f.err[dst] = f.err[src];
        for (int k = 0; k < 3; ++k) f.wc[k][dst] = f.wc[k][src];
        f.idx[dst] = f.idx[src];
        f.deg[dst] = f.deg[src];
        lists[dst].assign(f.nbr.begin() + f.nbr_off[src], f.nbr.begin() + f.nbr_off[src + 1]);
        newidx[src] = dst;
    }
    /* drop pairs touching deleted bubbles, remap moved indices, keep each
     * list ascending */
    std::vector<int32_t> off(nnew + 1, 0), lst;
    for (int32_t i = 0; i < nnew; ++i) {
        std::vector<int32_t> out;
        for (int32_t j : lists[i])
            if (newidx[j] >= 0) out.push_back(newidx[j]);
        std::sort(out.begin(), out.end());
        lst.insert(lst.end(), out.begin(), out.end());
        off[i + 1] = (int32_t)lst.size();
    }
    f.nbr_off.swap(off);
    f.nbr.swap(lst);
    for (int k = 0; k < 4; ++k) {
        f.x[k].resize(nnew);
        f.v[k].resize(nnew);
        f.vo[k].resize(nnew);
        f.saved[k].resize(nnew);
    }
    f.path.resize(nnew);
    f.err.resize(nnew);
    for (int k = 0; k < 3; ++k) f.wc[k].resize(nnew);
    f.idx.resize(nnew);
    f.deg.resize(nnew);
    /* volume fix: r *= M / S_V^(1/D), src/kernels.cu:861-867, :932-934 */
    double vm = 3 == c.dim ? 1.0 / det_cbrt(f.s.s_v) : 1.0 / std::sqrt(f.s.s_v);
    vm *= c.vol_mul;
    for (int32_t i = 0; i < nnew; ++i) f.x[3][i] = f.x[3][i] * vm;
    f.n = nnew;
    f.s.count = nnew;
    f.s.num_pairs = f.nbr_off[nnew] / 2;
    f.s.removals += 1;
}

/* integrate(): src/simulator.cu:251-337 */
int integrate(Foam &f, bool gas, bool incr_path, bool flow) {
    int tries = 0;
    do {
        attempt(f, gas, incr_path, flow);
        ++tries;
    } while (!f.s.accepted);
    const int need_search = f.s.need_search;
    commit(f, gas, incr_path);
    if (gas) {
        f.s.max_r_cell = f.s.max_r;
        if (f.s.n_del > 0) remove_bubbles(f);
    }
    if (need_search) search(f);
    return tries;
}

/* SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) seeding xoshiro256**
 * (Blackman and Vigna, 2018): the project RNG for synthetic initial foams. */
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int k = 0; k < 4; ++k) s[k] = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    /* Marsaglia polar method (Marsaglia and Bray, 1964), both variates used
     * in turn; needs no trigonometric function */
    bool have = false;
    double spare = 0.0;
    double normal() {
        if (have) {
            have = false;
            return spare;
        }
        double u, v, s;
        do {
            u = 2.0 * uniform() - 1.0;
            v = 2.0 * uniform() - 1.0;
            s = u * u + v * v;
        } while (s >= 1.0 || s == 0.0);
        const double m = std::sqrt(-2.0 * std::log(s) / s);
        spare = v * m;
        have = true;
        return u * m;
    }
};

Foam *new_foam(const ora_consts &c, int32_t n) {
    Foam *f = new Foam();
    f->c = c;
    f->n = n;
    for (int k = 0; k < 4; ++k) {
        f->x[k].assign(n, 0.0);
        f->v[k].assign(n, 0.0);
        f->vo[k].assign(n, 0.0);
        f->saved[k].assign(n, 0.0);
    }
    f->path.assign(n, 0.0);
    f->err.assign(n, 0.0);
    for (int k = 0; k < 3; ++k) f->wc[k].assign(n, 0);
    f->idx.assign(n, 0);
    f->nbr_off.assign(n + 1, 0);
    f->deg.assign(n, 0);
    f->s.count = n;
    return f;
}

}  // namespace

extern "C" {

void *ora_create(const ora_consts *c, int32_t count) { return new_foam(*c, count); }

void ora_destroy(void *p) { delete static_cast<Foam *>(p); }

void ora_get_consts(void *p, ora_consts *c) { *c = static_cast<Foam *>(p)->c; }

void ora_set_state(void *p, const ora_state *st, const ora_scalars *sc) {
    Foam &f = *static_cast<Foam *>(p);
    const int32_t n = sc->count;
    f.s = *sc;
    f.n = n;
    auto cp = [n](std::vector<double> &dst, const double *src) {
        dst.assign(n, 0.0);
        if (src) std::memcpy(dst.data(), src, sizeof(double) * n);
    };
    for (int k = 0; k < 4; ++k) {
        cp(f.x[k], st->x[k]);
        cp(f.v[k], st->v[k]);
        cp(f.vo[k], st->vo[k]);
        cp(f.saved[k], st->saved[k]);
    }
    cp(f.path, st->path);
    cp(f.err, st->err);
    for (int k = 0; k < 3; ++k) {
        f.wc[k].assign(n, 0);
        if (st->wc[k]) std::memcpy(f.wc[k].data(), st->wc[k], sizeof(int32_t) * n);
    }
    f.idx.assign(n, 0);
    if (st->idx) std::memcpy(f.idx.data(), st->idx, sizeof(int32_t) * n);
    f.nbr_off.assign(n + 1, 0);
    if (st->nbr_off) std::memcpy(f.nbr_off.data(), st->nbr_off, sizeof(int32_t) * (n + 1));
    f.nbr.assign(f.nbr_off[n], 0);
    if (st->nbr && f.nbr_off[n] > 0) std::memcpy(f.nbr.data(), st->nbr, sizeof(int32_t) * f.nbr_off[n]);
    f.s.num_pairs = f.nbr_off[n] / 2;
    f.deg.assign(n, 0);
    for (int32_t i = 0; i < n; ++i)
        f.deg[i] = st->deg ? st->deg[i] : f.nbr_off[i + 1] - f.nbr_off[i];
}

int ora_get_state(void *p, ora_state *st, ora_scalars *sc) {
    Foam &f = *static_cast<Foam *>(p);
    const int32_t n = f.n;
    f.s.count = n;
    f.s.num_pairs = f.nbr_off[n] / 2;
    if (sc) *sc = f.s;
    if (!st) return 0;
    auto cp = [n](double *dst, const std::vector<double> &src) {
        if (dst) std::memcpy(dst, src.data(), sizeof(double) * n);
    };
    for (int k = 0; k < 4; ++k) {
        cp(st->x[k], f.x[k]);
        cp(st->v[k], f.v[k]);
        cp(st->vo[k], f.vo[k]);
        cp(st->saved[k], f.saved[k]);
    }
    cp(st->path, f.path);
    cp(st->err, f.err);
    for (int k = 0; k < 3; ++k)
        if (st->wc[k]) std::memcpy(st->wc[k], f.wc[k].data(), sizeof(int32_t) * n);
    if (st->idx) std::memcpy(st->idx, f.idx.data(), sizeof(int32_t) * n);
    if (st->nbr_off) std::memcpy(st->nbr_off, f.nbr_off.data(), sizeof(int32_t) * (n + 1));
    if (st->deg) std::memcpy(st->deg, f.deg.data(), sizeof(int32_t) * n);
    if (st->nbr) {
        if ((int64_t)f.nbr.size() > st->nbr_cap) return -1;
        std::memcpy(st->nbr, f.nbr.data(), sizeof(int32_t) * f.nbr.size());
    }
    return 0;
}

void ora_get_attempt(void *p, ora_attempt_out *o) {
    Foam &f = *static_cast<Foam *>(p);
    const size_t n = f.xp[0].size();
    for (int k = 0; k < 4; ++k) {
        if (o->xc[k]) std::memcpy(o->xc[k], f.xp[k].data(), sizeof(double) * n);
        if (o->vp[k]) std::memcpy(o->vp[k], f.vp[k].data(), sizeof(double) * n);
    }
    if (o->ov) std::memcpy(o->ov, f.ov.data(), sizeof(double) * n);
    if (o->path_new) std::memcpy(o->path_new, f.path_new.data(), sizeof(double) * n);
    if (o->del) std::memcpy(o->del, f.del.data(), n);
}

int ora_attempt(void *p, int gas, int path, int flow) {
    return attempt(*static_cast<Foam *>(p), gas != 0, path != 0, flow != 0);
}

int ora_integrate(void *p, int gas, int path, int flow) {
    return integrate(*static_cast<Foam *>(p), gas != 0, path != 0, flow != 0);
}

double ora_run(void *p, int gas, int path, int flow, int64_t n) {
    Foam &f = *static_cast<Foam *>(p);
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t k = 0; k < n; ++k) integrate(f, gas != 0, path != 0, flow != 0);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

void ora_search(void *p) { search(*static_cast<Foam *>(p)); }

void ora_remove(void *p) { remove_bubbles(*static_cast<Foam *>(p)); }

int64_t ora_brute_pairs(void *p, int32_t *out, int64_t cap) {
    Foam &f = *static_cast<Foam *>(p);
    const ora_consts &c = f.c;
    const int D = c.dim;
    int64_t np = 0;
    for (int32_t i = 0; i < f.n; ++i) {
        for (int32_t j = i + 1; j < f.n; ++j) {
            const double maxd = f.x[3][i] + f.x[3][j] + c.skin;
            double d[3];
            wrapped_diff(c, f.x[0][i], f.x[1][i], 3 == D ? f.x[2][i] : 0.0, f.x[0][j], f.x[1][j],
                         3 == D ? f.x[2][j] : 0.0, d);
            if (sq_len(d) < maxd * maxd) {
                if (np < cap) {
                    out[2 * np] = i;
                    out[2 * np + 1] = j;
                }
                ++np;
            }
        }
    }
    return np;
}

/* Synthetic initial foam.  Radii: clipped normal or log-normal; box edge
 * chosen so that sum V(r) / V_box = phi exactly; positions uniform in the
 * box.  Constants follow src/simulator.cu:546-575 and :820-826; the mean
 * initial area follows src/kernels.cu:1065-1069 with src/simulator.cu:708-716. */
void *ora_init_synthetic(const ora_init_params *p) {
    const int D = p->dim;
    const int32_t n = p->count;
    Rng rng(p->seed);
    std::vector<double> r(n);
    for (int32_t i = 0; i < n; ++i) {
        const double z = rng.normal();
        double v;
        if (1 == p->radius_dist) {
            v = p->r_mean * std::exp(p->r_std * z);
        } else if (2 == p->radius_dist) {
            /* src/kernels.cu:1051-1053: |r|, plus minRad when at or below it */
            const double min_rad = 0.1 * p->r_mean;
            v = std::fabs(p->r_mean + p->r_std * z);
            v = v > min_rad ? v : v + min_rad;
        } else {
            v = p->r_mean + p->r_std * z;
            v = v < p->r_lo ? p->r_lo : (v > p->r_hi ? p->r_hi : v);
        }
        r[i] = v;
    }
    /* total bubble volume: src/kernels.cu:1016-1020 */
    double vtot = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        double vol = r[i] * r[i] * kPi;
        if (3 == D) vol *= r[i] * 1.33333333333333333333333333;
        vtot += vol;
    }
    double rel = p->rel_dim[0] * p->rel_dim[1];
    if (3 == D) rel *= p->rel_dim[2];
    double t = vtot / (p->phi * rel);
    t = 3 == D ? std::cbrt(t) : std::sqrt(t);
    ora_consts c{};
    c.dim = D;
    for (int k = 0; k < 3; ++k) {
        c.wall[k] = k < D ? p->wall[k] : 0;
        c.lbb[k] = 0.0;
        c.tfr[k] = k < D ? t * p->rel_dim[k] : 0.0;
        c.interval[k] = c.tfr[k] - c.lbb[k];
    }
    c.wall_drag = p->wall_drag;
    c.skin = 0.3 * p->avg_rad;
    c.min_rad = 0.1 * p->avg_rad;
    c.f0 = (float)p->sigma * p->avg_rad / p->mu;
    c.k_param = p->k_param;
    c.kappa = p->kappa;
    c.tol = p->tol;
    double vbox = c.interval[0] * c.interval[1];
    if (3 == D) vbox *= c.interval[2];
    double mult = p->phi * vbox / kPi;
    c.vol_mul = 3 == D ? std::cbrt(0.75 * mult) : std::sqrt(mult);
    double ain = 0.0, maxr = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        double area = 2.0 * kPi * r[i];
        if (3 == D) area *= 2.0 * r[i];
        ain += area / n;
        maxr = std::fmax(maxr, r[i]);
    }
    c.avg_area_in = ain;
    Foam *f = new_foam(c, n);
    for (int32_t i = 0; i < n; ++i) {
        for (int k = 0; k < D; ++k) f->x[k][i] = c.lbb[k] + c.interval[k] * rng.uniform();
        f->x[3][i] = r[i];
        f->idx[i] = i;
    }
    f->s.dt = p->dt0;
    f->s.max_r_cell = maxr;
    /* cell capacity: factor times the initial cell count */
    f->c.cell_cap = 1 << 30;
    int32_t cd[3];
    cell_dims(*f, cd);
    double cap = p->cell_cap_factor * (double)cd[0] * cd[1] * cd[2];
    if (cap < 64.0) cap = 64.0;
    f->c.cell_cap = (int32_t)std::fmin(cap, 2.0e9);
    search(*f);
    f->s.searches = 0;
    return f;
}

/* reference src/simulator.cu:39-58 with src/kernels.cu:941-956 (totalEnergy,
 * potentialEnergy: for each pair, e = r_i + r_j - |wrappedDifference|, e^2
 * added to both bubbles when e > 0), src/simulator.cu:362-379 with
 * src/kernels.cu:1013-1024 (totalVolume, calculateVolumes) and the sums
 * behind getAvg (src/simulator.cu:1036-1050).  Each bubble sums its full
 * list in ascending neighbour order (the reference adds with atomics, in
 * unspecified order); bubbles are summed in index order. */
void ora_observe(void *p, double out[8], double *energy) {
    const Foam &f = *static_cast<Foam *>(p);
    const ora_consts &c = f.c;
    for (int q = 0; q < 8; ++q) out[q] = 0.0;
    for (int32_t i = 0; i < f.n; ++i) {
        const double zi = 3 == c.dim ? f.x[2][i] : 0.0;
        double e = 0.0;
        for (int32_t k = f.nbr_off[i]; k < f.nbr_off[i + 1]; ++k) {
            const int32_t j = f.nbr[k];
            double d[3];
            wrapped_diff(c, f.x[0][i], f.x[1][i], zi, f.x[0][j], f.x[1][j],
                         3 == c.dim ? f.x[2][j] : 0.0, d);
            const double o = f.x[3][i] + f.x[3][j] - std::sqrt(sq_len(d));
            if (o > 0) e += o * o;
        }
        if (energy) energy[i] = e;
        const double r = f.x[3][i];
        double vol = r * r * kPi;
        if (3 == c.dim) vol *= r * 1.33333333333333333333333333;
        out[0] += e;
        out[1] += vol;
        out[2] += r;
        out[3] += f.path[i];
        for (int q = 0; q < 4; ++q)
            if (2 != q || 3 == c.dim) out[4 + q] += f.v[q][i];
    }
}

/* The box rescale of init (reference src/simulator.cu:793-846, kernel
 * transformPositions src/kernels.cu:970-990): box t * rel_dim with t from
 * the bubbles' total volume `vol` and phi, coordinates and flow box kept
 * relative, the volume multiplier from the new box, then a search. */
void ora_rescale(void *p, double phi, const double rel[3], double vol) {
    Foam &f = *static_cast<Foam *>(p);
    ora_consts &c = f.c;
    const int D = c.dim;
    double t = vol / (phi * rel[0] * rel[1]);
    if (3 == D) {
        t /= rel[2];
        t = std::cbrt(t);
    } else {
        t = std::sqrt(t);
    }
    const ora_consts old = c;
    for (int k = 0; k < 3; ++k) {
        c.tfr[k] = k < D ? t * rel[k] : 0.0;
        c.interval[k] = c.tfr[k] - c.lbb[k];
        if (k < D) {
            c.flow_lbb[k] = c.interval[k] * ((old.flow_lbb[k] - old.lbb[k]) / old.interval[k]) + c.lbb[k];
            c.flow_tfr[k] = c.interval[k] * ((old.flow_tfr[k] - old.lbb[k]) / old.interval[k]) + c.lbb[k];
        }
    }
    for (int k = 0; k < D; ++k)
        for (int32_t i = 0; i < f.n; ++i) {
            const double r = (f.x[k][i] - old.lbb[k]) / old.interval[k];
            f.x[k][i] = c.interval[k] * r + c.lbb[k];
        }
    double vbox = c.interval[0] * c.interval[1];
    if (3 == D) vbox *= c.interval[2];
    const double mult = phi * vbox / kPi;
    c.vol_mul = 3 == D ? std::cbrt(0.75 * mult) : std::sqrt(mult);
    search(f);
}

/* The AB2 velocity bootstrap of init (reference src/simulator.cu:739-787,
 * kernel euler src/kernels.cu:958-968): with the predicted state equal to
 * the current one, the pair velocities v0 (no gas exchange, no flow, no
 * walls); an Euler step x1 = x0 + v0 dt of the predicted positions; the
 * pair velocities at x1 become the old velocities.  The current velocities
 * are left as they are (zero at creation). */
void ora_bootstrap(void *p) {
    Foam &f = *static_cast<Foam *>(p);
    const int D = f.c.dim;
    const double ts = f.s.dt;
    resize_attempt(f);
    for (int k = 0; k < 4; ++k) f.xp[k] = f.x[k];
    double s_oa = 0.0, s_oar = 0.0;
    pair_step(f, false, false, s_oa, s_oar);
    for (int k = 0; k < D; ++k)
        for (int32_t i = 0; i < f.n; ++i) f.xp[k][i] += f.vp[k][i] * ts;
    pair_step(f, false, false, s_oa, s_oar);
    for (int k = 0; k < D; ++k) f.vo[k] = f.vp[k];
}

double ora_cbrt(double a) { return det_cbrt(a); }

void ora_wall_velocity(const ora_consts *c, const double pos[3], double rad, double v[3]) {
    add_wall_velocity(*c, pos[0], pos[1], pos[2], rad, v[0], v[1], v[2]);
}

void ora_splitmix64(uint64_t seed, int32_t n, uint64_t *out) {
    uint64_t x = seed;
    for (int32_t k = 0; k < n; ++k) out[k] = Rng::splitmix(x);
}

void ora_xoshiro256ss(const uint64_t state[4], int32_t n, uint64_t *out) {
    Rng r(0);
    for (int k = 0; k < 4; ++k) r.s[k] = state[k];
    for (int32_t k = 0; k < n; ++k) out[k] = r.next();
}

void ora_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

int ora_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
