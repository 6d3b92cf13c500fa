// state_io.hip -- copies of one foam's state between the device and
// caller-owned host views (cubble_state_view, cubble_attempt_view,
// cubble_scalars).  Neighbour lists travel as CSR on the host side and as
// column-major fixed-capacity lists on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "engine.hpp"

namespace cubble_amd {

#define R_TRY(eng, call)                                                                \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return (eng).fail(e_ == hipErrorOutOfMemory ? CUBBLE_E_NOMEM : CUBBLE_E_HIP, \
                              std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

int get_scalars(Engine &e, int k, cubble_scalars *o) {
    if (!o || k < 0 || k >= (int)e.foams.size()) return e.fail(CUBBLE_E_INVALID, "bad argument");
    Foam f;
    int rc = e.read_header(k, f);
    if (rc) return rc;
    std::memset(o, 0, sizeof(*o));
    o->dim = f.c.dim;
    o->count = f.d.count;
    o->capacity = f.cap;
    o->nbr_cap = f.nbr_cap;
    o->cell_cap = f.cell_cap;
    for (int q = 0; q < 3; ++q) {
        o->cell_dim[q] = f.d.cell_dim[q];
        o->wall[q] = f.c.wall[q];
        o->lbb[q] = f.c.lbb[q];
        o->tfr[q] = f.c.tfr[q];
        o->interval[q] = f.c.interval[q];
    }
    o->precision = f.precision;
    o->wall_drag = f.c.wall_drag;
    o->skin = f.c.skin;
    o->min_rad = f.c.min_rad;
    o->f0 = f.c.f0;
    o->k_param = f.c.k_param;
    o->kappa = f.c.kappa;
    o->avg_area_in = f.c.avg_area_in;
    o->vol_mul = f.c.vol_mul;
    o->tol = f.c.tol;
    o->dt = f.d.dt;
    o->time_int = f.d.time_int;
    o->time_frac = f.d.time_frac;
    o->max_r_cell = f.d.max_r_cell;
    o->steps = f.d.steps;
    o->attempts = f.d.attempts;
    o->searches = f.d.searches;
    o->removals = f.d.removals;
    o->bubble_steps = f.d.bubble_steps;
    o->ts_min = f.d.ts_min;
    o->ts_max = f.d.ts_max;
    o->ts_sum = f.d.ts_sum;
    for (int q = 0; q < 3; ++q) {
        o->flow_lbb[q] = f.c.flow_lbb[q];
        o->flow_tfr[q] = f.c.flow_tfr[q];
        o->flow_vel[q] = f.c.flow_vel[q];
    }
    o->s_a = f.d.s_a;
    o->s_ar = f.d.s_ar;
    o->s_oa = f.d.s_oa;
    o->s_oar = f.d.s_oar;
    o->s_v = f.d.s_v;
    o->max_err = f.d.max_err;
    o->max_r = f.d.max_r;
    o->max_exp = f.d.max_exp;
    o->n_del = f.d.n_del;
    o->accepted = f.d.accepted;
    o->need_search = f.d.need_search;
    o->num_pairs = f.d.num_pairs;
    return CUBBLE_OK;
}

namespace {
int d2h(Engine &e, void *dst, const void *src, size_t bytes) {
    if (!dst || !bytes) return CUBBLE_OK;
    R_TRY(e, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return CUBBLE_OK;
}

int h2d(Engine &e, void *dst, const void *src, size_t bytes) {
    if (!src || !bytes) return CUBBLE_OK;
    R_TRY(e, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return CUBBLE_OK;
}

// Velocities, saved positions, error and overlap are FP32 in mixed precision,
// FP64 otherwise; host views are FP64 in both.
int d2h_field(Engine &e, const Foam &f, double *dst, const void *src, size_t n) {
    if (!dst || !n) return CUBBLE_OK;
    if (CUBBLE_PREC_MIXED != f.precision) return d2h(e, dst, src, n * sizeof(double));
    std::vector<float> t(n);
    int rc = d2h(e, t.data(), src, n * sizeof(float));
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) dst[i] = (double)t[i];
    return CUBBLE_OK;
}

int h2d_field(Engine &e, const Foam &f, void *dst, const double *src, size_t n) {
    if (!src || !n) return CUBBLE_OK;
    if (CUBBLE_PREC_MIXED != f.precision) return h2d(e, dst, src, n * sizeof(double));
    std::vector<float> t(n);
    for (size_t i = 0; i < n; ++i) t[i] = (float)src[i];
    return h2d(e, dst, t.data(), n * sizeof(float));
}

}  // namespace

int get_state(Engine &e, int k, cubble_state_view *v) {
    if (!v || k < 0 || k >= (int)e.foams.size()) return e.fail(CUBBLE_E_INVALID, "bad argument");
    Foam f;
    int rc = e.read_header(k, f);
    if (rc) return rc;
    const size_t n = (size_t)f.d.count;
    const size_t db = n * sizeof(double), ib = n * sizeof(int32_t);
    for (int q = 0; q < 4; ++q) {
        if (2 == q && 2 == f.c.dim) {
            for (double *p : {v->x[q], v->v[q], v->vo[q], v->saved[q]})
                if (p) std::memset(p, 0, db);
            continue;
        }
        if ((rc = d2h(e, v->x[q], f.r.x[q], db)) || (rc = d2h_field(e, f, v->v[q], f.r.v[q], n)) ||
            (rc = d2h_field(e, f, v->vo[q], f.r.vo[q], n)) ||
            (rc = d2h_field(e, f, v->saved[q], f.r.saved[q], n)))
            return rc;
    }
    if ((rc = d2h(e, v->path, f.r.path, db)) || (rc = d2h_field(e, f, v->error, f.r.err, n)))
        return rc;
    for (int q = 0; q < 3; ++q) {
        if (q >= f.c.dim) {
            if (v->wrap[q]) std::memset(v->wrap[q], 0, ib);
            continue;
        }
        if ((rc = d2h(e, v->wrap[q], f.r.wc[q], ib))) return rc;
    }
    if ((rc = d2h(e, v->index, f.r.idx, ib)) || (rc = d2h(e, v->degree, f.s.deg, ib))) return rc;
    if (v->nbr_off || v->nbr) {
        std::vector<int32_t> nn(n);
        if ((rc = d2h(e, nn.data(), f.s.nn, ib))) return rc;
        int maxn = 0;
        for (size_t i = 0; i < n; ++i) maxn = std::max(maxn, nn[i]);
        std::vector<int32_t> ell((size_t)maxn * n);
        if (maxn > 0)
            R_TRY(e, hipMemcpy2D(ell.data(), n * sizeof(int32_t), f.s.nbr,
                                 (size_t)f.cap * sizeof(int32_t), n * sizeof(int32_t), maxn,
                                 hipMemcpyDeviceToHost));
        int64_t tot = 0;
        if (v->nbr_off) v->nbr_off[0] = 0;
        for (size_t i = 0; i < n; ++i) {
            if (v->nbr) {
                if (tot + nn[i] > v->nbr_cap)
                    return e.fail(CUBBLE_E_CAPACITY, "get_state: nbr buffer too small");
                for (int q = 0; q < nn[i]; ++q) v->nbr[tot + q] = ell[(size_t)q * n + i];
            }
            tot += nn[i];
            if (v->nbr_off) v->nbr_off[i + 1] = (int32_t)tot;
        }
    }
    return CUBBLE_OK;
}

int get_attempt(Engine &e, int k, cubble_attempt_view *v) {
    if (!v || k < 0 || k >= (int)e.foams.size()) return e.fail(CUBBLE_E_INVALID, "bad argument");
    Foam f;
    int rc = e.read_header(k, f);
    if (rc) return rc;
    const size_t n = (size_t)f.d.count;
    const size_t db = n * sizeof(double);
    for (int q = 0; q < 4; ++q) {
        if (2 == q && 2 == f.c.dim) {
            if (v->xc[q]) std::memset(v->xc[q], 0, db);
            if (v->vp[q]) std::memset(v->vp[q], 0, db);
            continue;
        }
        if ((rc = d2h(e, v->xc[q], f.r.xp[q], db)) || (rc = d2h_field(e, f, v->vp[q], f.r.vp[q], n)))
            return rc;
    }
    if ((rc = d2h_field(e, f, v->overlap, f.s.ov, n)) ||
        (rc = d2h(e, v->path_new, f.r.path_new, db)) ||
        (rc = d2h(e, v->deleted, f.s.del, n)))
        return rc;
    return CUBBLE_OK;
}

namespace {
// CSR neighbour lists as the kernels need them: offsets from 0, non-
// decreasing, within the view's buffer; entries in [0, n), never the bubble
// itself, strictly ascending per list, and symmetric (j lists i whenever i
// lists j); the flow's degrees, if given, not negative.
int check_lists(Engine &e, const cubble_state_view *v, int n) {
    const int32_t *off = v->nbr_off;
    if (off[0] != 0) return e.fail(CUBBLE_E_INVALID, "set_state: nbr_off[0] must be 0");
    for (int i = 0; i < n; ++i)
        if (off[i + 1] < off[i])
            return e.fail(CUBBLE_E_INVALID, "set_state: nbr_off must not decrease");
    if (off[n] > 0 && (!v->nbr || (int64_t)off[n] > v->nbr_cap))
        return e.fail(CUBBLE_E_INVALID, "set_state: nbr_off[count] exceeds the nbr buffer");
    for (int i = 0; i < n; ++i) {
        for (int32_t q = off[i]; q < off[i + 1]; ++q) {
            const int32_t j = v->nbr[q];
            if (j < 0 || j >= n || j == i)
                return e.fail(CUBBLE_E_INVALID, "set_state: list entry " + std::to_string(j) +
                                                    " of bubble " + std::to_string(i) +
                                                    " is out of range or the bubble itself");
            if (q > off[i] && j <= v->nbr[q - 1])
                return e.fail(CUBBLE_E_INVALID, "set_state: list of bubble " + std::to_string(i) +
                                                    " is not strictly ascending");
        }
    }
    for (int i = 0; i < n; ++i)
        for (int32_t q = off[i]; q < off[i + 1]; ++q) {
            const int32_t j = v->nbr[q];
            if (!std::binary_search(v->nbr + off[j], v->nbr + off[j + 1], i))
                return e.fail(CUBBLE_E_INVALID, "set_state: lists are not symmetric (" +
                                                    std::to_string(i) + " lists " +
                                                    std::to_string(j) + ")");
        }
    if (v->degree)
        for (int i = 0; i < n; ++i)
            if (v->degree[i] < 0) return e.fail(CUBBLE_E_INVALID, "set_state: negative degree");
    return CUBBLE_OK;
}
}  // namespace

int set_state(Engine &e, int k, const cubble_state_view *v, const cubble_scalars *sc) {
    if (!v || !sc || k < 0 || k >= (int)e.foams.size())
        return e.fail(CUBBLE_E_INVALID, "bad argument");
    Foam f;
    int rc = e.read_header(k, f);
    if (rc) return rc;
    const int n = sc->count;
    if (n < 0 || n > f.cap) return e.fail(CUBBLE_E_CAPACITY, "set_state: count exceeds capacity");
    // the kernels index with the list entries: check every list before any
    // upload, so a rejected view leaves the state untouched
    if (v->nbr_off && (rc = check_lists(e, v, n))) return rc;
    const size_t db = (size_t)n * sizeof(double), ib = (size_t)n * sizeof(int32_t);
    for (int q = 0; q < 4; ++q) {
        if (2 == q && 2 == f.c.dim) continue;
        if ((rc = h2d(e, f.r.x[q], v->x[q], db)) || (rc = h2d_field(e, f, f.r.v[q], v->v[q], n)) ||
            (rc = h2d_field(e, f, f.r.vo[q], v->vo[q], n)) ||
            (rc = h2d_field(e, f, f.r.saved[q], v->saved[q], n)))
            return rc;
    }
    if ((rc = h2d(e, f.r.path, v->path, db)) || (rc = h2d_field(e, f, f.r.err, v->error, n)))
        return rc;
    for (int q = 0; q < f.c.dim; ++q)
        if ((rc = h2d(e, f.r.wc[q], v->wrap[q], ib))) return rc;
    if ((rc = h2d(e, f.r.idx, v->index, ib)) || (rc = h2d(e, f.s.deg, v->degree, ib))) return rc;
    if (v->nbr_off) {
        int maxn = 0;
        for (int i = 0; i < n; ++i) maxn = std::max(maxn, v->nbr_off[i + 1] - v->nbr_off[i]);
        if (maxn > f.nbr_cap) {
            if ((rc = e.grow(k, maxn))) return rc;
            if ((rc = e.read_header(k, f))) return rc;
        }
        std::vector<int32_t> nn(n), ell((size_t)std::max(maxn, 1) * n, 0);
        for (int i = 0; i < n; ++i) {
            nn[i] = v->nbr_off[i + 1] - v->nbr_off[i];
            for (int q = 0; q < nn[i]; ++q) ell[(size_t)q * n + i] = v->nbr[v->nbr_off[i] + q];
        }
        if ((rc = h2d(e, f.s.nn, nn.data(), ib))) return rc;
        if (!v->degree && (rc = h2d(e, f.s.deg, nn.data(), ib))) return rc;
        if (maxn > 0)
            R_TRY(e, hipMemcpy2D(f.s.nbr, (size_t)f.cap * sizeof(int32_t), ell.data(),
                                 (size_t)n * sizeof(int32_t), (size_t)n * sizeof(int32_t), maxn,
                                 hipMemcpyHostToDevice));
    }
    Dyn d = f.d;
    d.count = n;
    d.dt = sc->dt;
    d.time_int = sc->time_int;
    d.time_frac = sc->time_frac;
    d.max_r_cell = sc->max_r_cell;
    d.steps = sc->steps;
    d.attempts = sc->attempts;
    d.searches = sc->searches;
    d.removals = sc->removals;
    d.bubble_steps = sc->bubble_steps;
    d.num_pairs = v->nbr_off ? v->nbr_off[n] / 2 : d.num_pairs;
    for (int q = 0; q < 3; ++q) d.cell_dim[q] = sc->cell_dim[q];
    d.need_search = 0;
    d.need_remove = 0;
    d.halt = 0;
    d.overflow = 0;
    d.step_limit = 0;
    return e.write_dyn(k, d);
}


}  // namespace cubble_amd
