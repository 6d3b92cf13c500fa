// runtime.hip -- host runtime behind the C ABI (include/cubble_amd.h).
//
// A `cubble_sim` owns one foam: its device buffers, its device header
// (Foam), the block tables of its launches and a HIP stream.  The step
// control flow runs on the device (a one-block control kernel decides
// accept / reject, dt, removal and search), so a run of many steps is a
// queue of identical kernel groups, captured once into a HIP graph and
// replayed, with one host synchronisation per replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cubble_amd.h"
#include "foam.hpp"
#include "host_init.hpp"
#include "launch.hpp"

using namespace cubble_amd;

namespace {

std::string g_create_error;

const char *kVersion = "cubble_amd 0.1 (gfx950)";

}  // namespace

struct cubble_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    Foam h{};              // host mirror: constants, buffers, sizes
    Foam *dfoam = nullptr; // device header
    int2 *bt_bub = nullptr;
    int2 *bt_cell = nullptr;
    int2 *bt_scell = nullptr;
    std::vector<void *> allocs;
    Dyn *pinned = nullptr; // pinned readback of the dynamic scalars
    std::string err;
    int precision = CUBBLE_PREC_FP64;
    // graph of `graph_groups` integrate groups, keyed by launch flags
    hipGraphExec_t graph = nullptr;
    uint32_t graph_flags = 0;
    int graph_groups = 0;
    int num_end = 0;
};

namespace {

#define HIP_TRY(sim, call)                                                              \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (sim)->err = std::string(#call) + ": " + hipGetErrorString(e_);             \
            return e_ == hipErrorOutOfMemory ? CUBBLE_E_NOMEM : CUBBLE_E_HIP;           \
        }                                                                               \
    } while (0)

int fail(cubble_sim *s, int code, const std::string &m) {
    s->err = m;
    return code;
}

template <typename T>
int dalloc(cubble_sim *s, T **p, size_t count) {
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    HIP_TRY(s, hipMalloc(&q, bytes));
    HIP_TRY(s, hipMemset(q, 0, bytes));
    s->allocs.push_back(q);
    *p = static_cast<T *>(q);
    return CUBBLE_OK;
}

Grid grid(const cubble_sim *s) {
    Grid g;
    g.foams = s->dfoam;
    g.bt_bub = s->bt_bub;
    g.nb_bub = s->h.nblk;
    g.bt_cell = s->bt_cell;
    g.nb_cell = s->h.cblk;
    g.bt_scell = s->bt_scell;
    g.nb_scell = s->h.sblk;
    g.nfoams = 1;
    g.stream = s->stream;
    return g;
}

uint32_t flags_of(const cubble_integrate_params *ip) {
    uint32_t f = 0;
    if (!ip || ip->use_gas_exchange) f |= kGas;
    if (!ip || ip->increment_path) f |= kPath;
    return f;
}

int check_ip(cubble_sim *s, const cubble_integrate_params *ip) {
    if (ip && ip->use_flow) return fail(s, CUBBLE_E_UNSUPPORTED, "imposed flow is not supported");
    return CUBBLE_OK;
}

int upload_header(cubble_sim *s) {
    HIP_TRY(s, hipMemcpy(s->dfoam, &s->h, sizeof(Foam), hipMemcpyHostToDevice));
    return CUBBLE_OK;
}

int read_dyn(cubble_sim *s, Dyn &d) {
    HIP_TRY(s, hipMemcpyAsync(s->pinned, &s->dfoam->d, sizeof(Dyn), hipMemcpyDeviceToHost,
                              s->stream));
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    d = *s->pinned;
    return CUBBLE_OK;
}

int read_header(cubble_sim *s, Foam &f) {
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemcpy(&f, s->dfoam, sizeof(Foam), hipMemcpyDeviceToHost));
    return CUBBLE_OK;
}

int write_dyn(cubble_sim *s, const Dyn &d) {
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    HIP_TRY(s, hipMemcpy(&s->dfoam->d, &d, sizeof(Dyn), hipMemcpyHostToDevice));
    return CUBBLE_OK;
}

void fill_result(const Dyn &d, cubble_step_result *r) {
    if (!r) return;
    r->error_too_large = d.accepted ? 0 : 1;
    r->num_to_be_deleted = d.n_del;
    r->count = d.count;
    r->num_pairs = d.num_pairs;
    r->need_search = d.need_search;
    r->halted = d.halt;
    r->max_error = d.max_err;
    r->max_radius = d.max_r;
    r->max_expansion = d.max_exp;
    r->dt = d.dt;
    r->time = (double)d.time_int + d.time_frac;
    r->steps = d.steps;
    r->attempts = d.attempts;
    r->searches = d.searches;
    r->removals = d.removals;
    r->bubble_steps = d.bubble_steps;
}

// Raise the neighbour-list capacity to hold `need` entries per bubble;
// with `rebuild`, recompute the lists (after a search overflowed).
int grow_lists(cubble_sim *s, int need, bool rebuild = true) {
    Foam f;
    int rc = read_header(s, f);
    if (rc) return rc;
    int ncap = std::max(need + need / 4 + 4, s->h.nbr_cap * 2);
    int32_t *nbr = nullptr;
    rc = dalloc(s, &nbr, (size_t)ncap * s->h.cap);
    if (rc) return rc;
    // free the old list
    for (auto it = s->allocs.begin(); it != s->allocs.end(); ++it)
        if (*it == (void *)s->h.s.nbr) {
            (void)hipFree(*it);
            s->allocs.erase(it);
            break;
        }
    s->h.s.nbr = nbr;
    s->h.nbr_cap = ncap;
    HIP_TRY(s, hipMemcpy(&s->dfoam->s.nbr, &nbr, sizeof(nbr), hipMemcpyHostToDevice));
    HIP_TRY(s, hipMemcpy(&s->dfoam->nbr_cap, &ncap, sizeof(ncap), hipMemcpyHostToDevice));
    if (!rebuild) return CUBBLE_OK;
    const Grid g = grid(s);
    launch_set_flags(g, 1, -1, 1);
    launch_search_lists(g);
    HIP_TRY(s, hipGetLastError());
    Dyn d;
    rc = read_dyn(s, d);
    if (rc) return rc;
    if (d.overflow) return fail(s, CUBBLE_E_CAPACITY, "neighbour list capacity growth failed");
    // the rebuilt search counted once more; undo the duplicate count
    d.searches -= 1;
    return write_dyn(s, d);
}

// After device work: if a search overflowed the list capacity, grow and
// rebuild (the step itself was already committed).
int settle(cubble_sim *s, Dyn &d) {
    int rc = read_dyn(s, d);
    if (rc) return rc;
    if (d.overflow) {
        rc = grow_lists(s, d.max_nn);
        if (rc) return rc;
        rc = read_dyn(s, d);
        if (rc) return rc;
    }
    return CUBBLE_OK;
}

void launch_pair_any(cubble_sim *s, const Grid &g, uint32_t flags) {
    if (s->precision == CUBBLE_PREC_MIXED) launch_pair_mixed(g, flags);
    else launch_pair(g, flags);
}

// One step() attempt: predictor, pair, sums, corrector, control.
void enqueue_attempt(cubble_sim *s, uint32_t flags) {
    const Grid g = grid(s);
    launch_predict(g, flags);
    launch_pair_any(s, g, flags);
    launch_sums(g);
    launch_correct(g, flags);
}

// One integrate() group: an attempt, then removal and search, each of which
// runs only if the attempt was accepted and the device control asked for it.
void enqueue_group(cubble_sim *s, uint32_t flags) {
    const Grid g = grid(s);
    enqueue_attempt(s, flags);
    launch_remove(g);
    launch_search(g);
}

int build_graph(cubble_sim *s, uint32_t flags, int groups) {
    if (s->graph && s->graph_flags == flags && s->graph_groups == groups) return CUBBLE_OK;
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    hipGraph_t gr = nullptr;
    HIP_TRY(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < groups; ++k) enqueue_group(s, flags);
    HIP_TRY(s, hipStreamEndCapture(s->stream, &gr));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return fail(s, CUBBLE_E_HIP, std::string("hipGraphInstantiate: ") +
                                                      hipGetErrorString(e));
    s->graph = ex;
    s->graph_flags = flags;
    s->graph_groups = groups;
    return CUBBLE_OK;
}

int make_sim(const cubble_config *cfg, cubble_sim **out) {
    *out = nullptr;
    cubble_sim *s = new cubble_sim();
    std::vector<double> x4;
    InitOut io;
    std::string err;
    int rc = generate_foam(*cfg, x4, io, err);
    if (rc) {
        g_create_error = err;
        delete s;
        return rc;
    }
    auto bail = [&](int code) {
        g_create_error = s->err;
        cubble_destroy(s);
        return code;
    };
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        s->err = "no HIP device found";
        return bail(CUBBLE_E_HIP);
    }
    if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        s->err = "hipStreamCreate failed";
        return bail(CUBBLE_E_HIP);
    }
    const int n = cfg->num_start;
    Foam &F = s->h;
    std::memset(&F, 0, sizeof(F));
    F.c = io.c;
    F.cap = n;
    F.precision = cfg->precision;
    s->precision = cfg->precision;
    s->num_end = cfg->num_end;
    F.nbr_cap = cfg->nbr_cap > 0 ? cfg->nbr_cap : (3 == cfg->dim ? 48 : 24);
    F.cell_cap = io.cell_cap;
    F.ipt = std::max(1, (n + kBlock * kMaxBlocks - 1) / (kBlock * kMaxBlocks));
    F.nblk = (n + kBlock * F.ipt - 1) / (kBlock * F.ipt);
    F.cipt = std::max(1, (F.cell_cap + kBlock * 1024 - 1) / (kBlock * 1024));
    F.cblk = std::max(1, (F.cell_cap + kBlock * F.cipt - 1) / (kBlock * F.cipt));
    F.sblk = std::max(1, (F.cell_cap + kSearchCells - 1) / kSearchCells);
    const size_t cap = (size_t)n;
    // buffers: positions x2, velocities x4, saved, path x2, err x2
    double *xb[2][4], *vb[4][4];
    for (int a = 0; a < 2; ++a)
        for (int q = 0; q < 4; ++q)
            if ((rc = dalloc(s, &xb[a][q], cap))) return bail(rc);
    for (int a = 0; a < 4; ++a)
        for (int q = 0; q < 4; ++q)
            if ((rc = dalloc(s, &vb[a][q], cap))) return bail(rc);
    for (int q = 0; q < 4; ++q) {
        F.r.x[q] = xb[0][q];
        F.r.xp[q] = xb[1][q];
        F.r.v[q] = vb[0][q];
        F.r.vp[q] = vb[1][q];
        F.r.vo[q] = vb[2][q];
        F.r.vs[q] = vb[3][q];
        if ((rc = dalloc(s, &F.r.saved[q], cap))) return bail(rc);
    }
    if ((rc = dalloc(s, &F.r.path, cap)) || (rc = dalloc(s, &F.r.path_new, cap)) ||
        (rc = dalloc(s, &F.r.err, cap)) || (rc = dalloc(s, &F.r.err_alt, cap)))
        return bail(rc);
    for (int q = 0; q < 3; ++q)
        if ((rc = dalloc(s, &F.r.wc[q], cap)) || (rc = dalloc(s, &F.r.wc_alt[q], cap)))
            return bail(rc);
    if ((rc = dalloc(s, &F.r.idx, cap)) || (rc = dalloc(s, &F.r.idx_alt, cap))) return bail(rc);
    Scratch &S = F.s;
    if ((rc = dalloc(s, &S.p4, cap)) || (rc = dalloc(s, &S.x4, cap)) ||
        (rc = dalloc(s, &S.ov, cap)) || (rc = dalloc(s, &S.del, cap)) ||
        (rc = dalloc(s, &S.nn, cap)) || (rc = dalloc(s, &S.nbr, cap * F.nbr_cap)) ||
        (rc = dalloc(s, &S.cell_of, cap)) ||
        (rc = dalloc(s, &S.cpart_i, (size_t)F.sblk * kPartI)) ||
        (rc = dalloc(s, &S.cell_count, (size_t)F.cell_cap)) ||
        (rc = dalloc(s, &S.cell_start, (size_t)F.cell_cap + 1)) ||
        (rc = dalloc(s, &S.cell_cursor, (size_t)F.cell_cap)) ||
        (rc = dalloc(s, &S.members, cap)) ||
        (rc = dalloc(s, &S.part, (size_t)F.nblk * kPartD)) ||
        (rc = dalloc(s, &S.part_i, (size_t)F.nblk * kPartI)) ||
        (rc = dalloc(s, &S.cpart, (size_t)F.cblk)) || (rc = dalloc(s, &S.del_list, cap)) ||
        (rc = dalloc(s, &S.plan_src, cap)) || (rc = dalloc(s, &S.moved_to, cap)))
        return bail(rc);
    // initial state
    for (int q = 0; q < 4; ++q) {
        if (2 == q && 2 == cfg->dim) continue;
        if (hipMemcpy(F.r.x[q], x4.data() + (size_t)q * n, sizeof(double) * n,
                      hipMemcpyHostToDevice) != hipSuccess) {
            s->err = "upload of initial state failed";
            return bail(CUBBLE_E_HIP);
        }
    }
    std::vector<int32_t> iota(n);
    for (int i = 0; i < n; ++i) iota[i] = i;
    if (hipMemcpy(F.r.idx, iota.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess) {
        s->err = "upload of indices failed";
        return bail(CUBBLE_E_HIP);
    }
    Dyn &d = F.d;
    d.count = n;
    d.dt = cfg->dt0;
    d.max_r_cell = io.max_r;
    d.need_search = 1;
    // block tables
    std::vector<int2> bt(F.nblk), bc(F.cblk), bs(F.sblk);
    for (int b = 0; b < F.nblk; ++b) bt[b] = make_int2(0, b);
    for (int b = 0; b < F.cblk; ++b) bc[b] = make_int2(0, b);
    for (int b = 0; b < F.sblk; ++b) bs[b] = make_int2(0, b);
    if ((rc = dalloc(s, &s->bt_bub, bt.size())) || (rc = dalloc(s, &s->bt_cell, bc.size())) ||
        (rc = dalloc(s, &s->bt_scell, bs.size())))
        return bail(rc);
    if (hipMemcpy(s->bt_bub, bt.data(), sizeof(int2) * bt.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(s->bt_cell, bc.data(), sizeof(int2) * bc.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(s->bt_scell, bs.data(), sizeof(int2) * bs.size(), hipMemcpyHostToDevice) !=
            hipSuccess) {
        s->err = "upload of block tables failed";
        return bail(CUBBLE_E_HIP);
    }
    if ((rc = dalloc(s, &s->dfoam, 1))) return bail(rc);
    if (hipHostMalloc((void **)&s->pinned, sizeof(Dyn), hipHostMallocDefault) != hipSuccess) {
        s->err = "hipHostMalloc failed";
        return bail(CUBBLE_E_NOMEM);
    }
    if ((rc = upload_header(s))) return bail(rc);
    // first neighbour search (reference init, src/simulator.cu:725)
    launch_search(grid(s));
    if (hipGetLastError() != hipSuccess) {
        s->err = "launch of the first search failed";
        return bail(CUBBLE_E_HIP);
    }
    Dyn dn;
    if ((rc = settle(s, dn))) return bail(rc);
    dn.searches = 0;
    if ((rc = write_dyn(s, dn))) return bail(rc);
    *out = s;
    return CUBBLE_OK;
}

int copy_d2h(cubble_sim *s, void *dst, const void *src, size_t bytes) {
    if (!dst || !bytes) return CUBBLE_OK;
    HIP_TRY(s, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return CUBBLE_OK;
}

int copy_h2d(cubble_sim *s, void *dst, const void *src, size_t bytes) {
    if (!src || !bytes) return CUBBLE_OK;
    HIP_TRY(s, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return CUBBLE_OK;
}

}  // namespace

extern "C" {

const char *cubble_version(void) { return kVersion; }

void cubble_config_default(cubble_config *cfg) {
    if (cfg) config_default(*cfg);
}

int cubble_config_from_json(const char *json_path, cubble_config *out) {
    if (!out) return CUBBLE_E_INVALID;
    std::string err;
    const int rc = parse_config_json(json_path, *out, err);
    if (rc) g_create_error = err;
    return rc;
}

int cubble_create(const cubble_config *cfg, cubble_sim **out) {
    if (!cfg || !out) {
        g_create_error = "null argument";
        return CUBBLE_E_INVALID;
    }
    return make_sim(cfg, out);
}

int cubble_create_from_json(const char *json_path, cubble_sim **out) {
    if (!out) return CUBBLE_E_INVALID;
    cubble_config cfg;
    std::string err;
    const int rc = parse_config_json(json_path, cfg, err);
    if (rc) {
        g_create_error = err;
        *out = nullptr;
        return rc;
    }
    return make_sim(&cfg, out);
}

void cubble_destroy(cubble_sim *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    for (void *p : s->allocs) (void)hipFree(p);
    if (s->pinned) (void)hipHostFree(s->pinned);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

const char *cubble_last_error(const cubble_sim *s) {
    return s ? s->err.c_str() : g_create_error.c_str();
}

int cubble_generate_foam(const cubble_config *cfg, double *x4, cubble_scalars *sc) {
    if (!cfg) return CUBBLE_E_INVALID;
    std::vector<double> v;
    InitOut io;
    std::string err;
    const int rc = generate_foam(*cfg, v, io, err);
    if (rc) {
        g_create_error = err;
        return rc;
    }
    if (x4) std::memcpy(x4, v.data(), sizeof(double) * v.size());
    if (sc) {
        std::memset(sc, 0, sizeof(*sc));
        sc->dim = io.c.dim;
        sc->count = cfg->num_start;
        sc->capacity = cfg->num_start;
        sc->cell_cap = io.cell_cap;
        for (int k = 0; k < 3; ++k) {
            sc->wall[k] = io.c.wall[k];
            sc->lbb[k] = io.c.lbb[k];
            sc->tfr[k] = io.c.tfr[k];
            sc->interval[k] = io.c.interval[k];
        }
        sc->wall_drag = io.c.wall_drag;
        sc->skin = io.c.skin;
        sc->min_rad = io.c.min_rad;
        sc->f0 = io.c.f0;
        sc->k_param = io.c.k_param;
        sc->kappa = io.c.kappa;
        sc->avg_area_in = io.c.avg_area_in;
        sc->vol_mul = io.c.vol_mul;
        sc->tol = io.c.tol;
        sc->dt = cfg->dt0;
        sc->max_r_cell = io.max_r;
        sc->precision = cfg->precision;
    }
    return CUBBLE_OK;
}

int cubble_step(cubble_sim *s, const cubble_integrate_params *ip, cubble_step_result *res) {
    if (!s) return CUBBLE_E_INVALID;
    int rc = check_ip(s, ip);
    if (rc) return rc;
    const uint32_t flags = flags_of(ip);
    const Grid g = grid(s);
    launch_set_flags(g, -1, -1, 1);
    enqueue_attempt(s, flags);
    HIP_TRY(s, hipGetLastError());
    Dyn d;
    rc = read_dyn(s, d);
    if (rc) return rc;
    fill_result(d, res);
    return CUBBLE_OK;
}

int cubble_integrate(cubble_sim *s, const cubble_integrate_params *ip, cubble_step_result *res) {
    if (!s) return CUBBLE_E_INVALID;
    int rc = check_ip(s, ip);
    if (rc) return rc;
    const uint32_t flags = flags_of(ip) | kCommit;
    const Grid g = grid(s);
    launch_set_flags(g, -1, -1, 1);
    Dyn d{};
    for (int tries = 0; tries < 10000; ++tries) {
        enqueue_group(s, flags);
        HIP_TRY(s, hipGetLastError());
        rc = settle(s, d);
        if (rc) return rc;
        if (d.accepted) {
            fill_result(d, res);
            return CUBBLE_OK;
        }
    }
    return fail(s, CUBBLE_E_STATE, "integrate: no accepted step after 10000 attempts");
}

int cubble_run_steps(cubble_sim *s, const cubble_integrate_params *ip, int64_t num_steps,
                     cubble_step_result *res) {
    if (!s || num_steps < 0) return CUBBLE_E_INVALID;
    int rc = check_ip(s, ip);
    if (rc) return rc;
    const uint32_t flags = flags_of(ip) | kCommit;
    Dyn d;
    rc = read_dyn(s, d);
    if (rc) return rc;
    const int64_t target = d.steps + num_steps;
    d.step_limit = target;
    d.halt = 0;
    d.overflow = 0;
    rc = write_dyn(s, d);
    if (rc) return rc;
    const int groups = 16;
    rc = build_graph(s, flags, groups);
    if (rc) return rc;
    int64_t guard = 0;
    while (d.steps < target) {
        const int64_t left = target - d.steps;
        const int64_t launches = std::max<int64_t>(1, (left + groups - 1) / groups);
        for (int64_t k = 0; k < launches; ++k) HIP_TRY(s, hipGraphLaunch(s->graph, s->stream));
        rc = settle(s, d);
        if (rc) return rc;
        if (d.halt && d.steps < target) {
            // halted for a capacity rebuild (settle handled it): resume
            d.halt = 0;
            rc = write_dyn(s, d);
            if (rc) return rc;
        }
        if (++guard > 1000000) return fail(s, CUBBLE_E_STATE, "run_steps made no progress");
    }
    d.step_limit = 0;
    d.halt = 0;
    rc = write_dyn(s, d);
    if (rc) return rc;
    fill_result(d, res);
    return CUBBLE_OK;
}

int cubble_time_kernels(cubble_sim *s, const cubble_integrate_params *ip, int32_t reps,
                        cubble_kernel_times *out) {
    if (!s || !out || reps < 1) return CUBBLE_E_INVALID;
    int rc = check_ip(s, ip);
    if (rc) return rc;
    const uint32_t flags = flags_of(ip);
    const Grid g = grid(s);
    Dyn saved;
    rc = read_dyn(s, saved);
    if (rc) return rc;
    launch_set_flags(g, -1, -1, 1);
    std::vector<hipEvent_t> ev((size_t)reps * 4);
    for (auto &e : ev) HIP_TRY(s, hipEventCreate(&e));
    for (int r = 0; r < reps; ++r) {
        hipEvent_t *e = &ev[(size_t)r * 4];
        HIP_TRY(s, hipEventRecord(e[0], s->stream));
        launch_predict(g, flags);
        HIP_TRY(s, hipEventRecord(e[1], s->stream));
        launch_pair_any(s, g, flags);
        HIP_TRY(s, hipEventRecord(e[2], s->stream));
        // the corrector bucket includes the two one-block reduction kernels
        launch_sums(g);
        launch_correct(g, flags);
        HIP_TRY(s, hipEventRecord(e[3], s->stream));
    }
    HIP_TRY(s, hipStreamSynchronize(s->stream));
    double t[3] = {0.0, 0.0, 0.0};
    for (int r = 0; r < reps; ++r) {
        hipEvent_t *e = &ev[(size_t)r * 4];
        for (int k = 0; k < 3; ++k) {
            float ms = 0.f;
            HIP_TRY(s, hipEventElapsedTime(&ms, e[k], e[k + 1]));
            t[k] += ms;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    out->predict_ms = t[0] / reps;
    out->pair_ms = t[1] / reps;
    out->correct_ms = t[2] / reps;
    out->reps = reps;
    // the attempts above only changed dt and counters: restore them
    Dyn now;
    rc = read_dyn(s, now);
    if (rc) return rc;
    now.dt = saved.dt;
    now.attempts = saved.attempts;
    return write_dyn(s, now);
}

int cubble_search_neighbors(cubble_sim *s) {
    if (!s) return CUBBLE_E_INVALID;
    const Grid g = grid(s);
    launch_set_flags(g, 1, -1, 1);
    launch_search(g);
    HIP_TRY(s, hipGetLastError());
    Dyn d;
    return settle(s, d);
}

int cubble_remove_bubbles(cubble_sim *s) {
    if (!s) return CUBBLE_E_INVALID;
    const Grid g = grid(s);
    launch_set_flags(g, -1, 1, 0);
    launch_remove(g);
    HIP_TRY(s, hipGetLastError());
    Dyn d;
    return read_dyn(s, d);
}

int cubble_get_scalars(cubble_sim *s, cubble_scalars *o) {
    if (!s || !o) return CUBBLE_E_INVALID;
    Foam f;
    int rc = read_header(s, f);
    if (rc) return rc;
    std::memset(o, 0, sizeof(*o));
    o->dim = f.c.dim;
    o->count = f.d.count;
    o->capacity = f.cap;
    o->nbr_cap = f.nbr_cap;
    o->cell_cap = f.cell_cap;
    for (int k = 0; k < 3; ++k) {
        o->cell_dim[k] = f.d.cell_dim[k];
        o->wall[k] = f.c.wall[k];
        o->lbb[k] = f.c.lbb[k];
        o->tfr[k] = f.c.tfr[k];
        o->interval[k] = f.c.interval[k];
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

int cubble_get_state(cubble_sim *s, cubble_state_view *v) {
    if (!s || !v) return CUBBLE_E_INVALID;
    Foam f;
    int rc = read_header(s, f);
    if (rc) return rc;
    const size_t n = (size_t)f.d.count;
    const size_t db = n * sizeof(double), ib = n * sizeof(int32_t);
    for (int q = 0; q < 4; ++q) {
        if (2 == q && 2 == f.c.dim) {
            for (double *p : {v->x[q], v->v[q], v->vo[q], v->saved[q]})
                if (p) std::memset(p, 0, db);
            continue;
        }
        if ((rc = copy_d2h(s, v->x[q], f.r.x[q], db)) || (rc = copy_d2h(s, v->v[q], f.r.v[q], db)) ||
            (rc = copy_d2h(s, v->vo[q], f.r.vo[q], db)) ||
            (rc = copy_d2h(s, v->saved[q], f.r.saved[q], db)))
            return rc;
    }
    if ((rc = copy_d2h(s, v->path, f.r.path, db)) || (rc = copy_d2h(s, v->error, f.r.err, db)))
        return rc;
    for (int q = 0; q < 3; ++q) {
        if (q >= f.c.dim) {
            if (v->wrap[q]) std::memset(v->wrap[q], 0, ib);
            continue;
        }
        if ((rc = copy_d2h(s, v->wrap[q], f.r.wc[q], ib))) return rc;
    }
    if ((rc = copy_d2h(s, v->index, f.r.idx, ib))) return rc;
    if (v->nbr_off || v->nbr) {
        std::vector<int32_t> nn(n);
        if ((rc = copy_d2h(s, nn.data(), f.s.nn, ib))) return rc;
        int maxn = 0;
        for (size_t i = 0; i < n; ++i) maxn = std::max(maxn, nn[i]);
        std::vector<int32_t> ell((size_t)maxn * n);
        if (maxn > 0) {
            HIP_TRY(s, hipMemcpy2D(ell.data(), n * sizeof(int32_t), f.s.nbr,
                                   (size_t)f.cap * sizeof(int32_t), n * sizeof(int32_t), maxn,
                                   hipMemcpyDeviceToHost));
        }
        int64_t tot = 0;
        if (v->nbr_off) v->nbr_off[0] = 0;
        for (size_t i = 0; i < n; ++i) {
            if (v->nbr) {
                if (tot + nn[i] > v->nbr_cap)
                    return fail(s, CUBBLE_E_CAPACITY, "get_state: nbr buffer too small");
                for (int q = 0; q < nn[i]; ++q) v->nbr[tot + q] = ell[(size_t)q * n + i];
            }
            tot += nn[i];
            if (v->nbr_off) v->nbr_off[i + 1] = (int32_t)tot;
        }
    }
    return CUBBLE_OK;
}

int cubble_get_attempt(cubble_sim *s, cubble_attempt_view *v) {
    if (!s || !v) return CUBBLE_E_INVALID;
    Foam f;
    int rc = read_header(s, f);
    if (rc) return rc;
    const size_t n = (size_t)f.d.count;
    const size_t db = n * sizeof(double);
    for (int q = 0; q < 4; ++q) {
        if (2 == q && 2 == f.c.dim) {
            if (v->xc[q]) std::memset(v->xc[q], 0, db);
            if (v->vp[q]) std::memset(v->vp[q], 0, db);
            continue;
        }
        if ((rc = copy_d2h(s, v->xc[q], f.r.xp[q], db)) ||
            (rc = copy_d2h(s, v->vp[q], f.r.vp[q], db)))
            return rc;
    }
    if ((rc = copy_d2h(s, v->overlap, f.s.ov, db)) ||
        (rc = copy_d2h(s, v->path_new, f.r.path_new, db)) ||
        (rc = copy_d2h(s, v->deleted, f.s.del, n)))
        return rc;
    return CUBBLE_OK;
}

int cubble_set_state(cubble_sim *s, const cubble_state_view *v, const cubble_scalars *sc) {
    if (!s || !v || !sc) return CUBBLE_E_INVALID;
    Foam f;
    int rc = read_header(s, f);
    if (rc) return rc;
    const int n = sc->count;
    if (n < 0 || n > f.cap) return fail(s, CUBBLE_E_CAPACITY, "set_state: count exceeds capacity");
    const size_t db = (size_t)n * sizeof(double), ib = (size_t)n * sizeof(int32_t);
    for (int q = 0; q < 4; ++q) {
        if (2 == q && 2 == f.c.dim) continue;
        if ((rc = copy_h2d(s, f.r.x[q], v->x[q], db)) || (rc = copy_h2d(s, f.r.v[q], v->v[q], db)) ||
            (rc = copy_h2d(s, f.r.vo[q], v->vo[q], db)) ||
            (rc = copy_h2d(s, f.r.saved[q], v->saved[q], db)))
            return rc;
    }
    if ((rc = copy_h2d(s, f.r.path, v->path, db)) || (rc = copy_h2d(s, f.r.err, v->error, db)))
        return rc;
    for (int q = 0; q < f.c.dim; ++q)
        if ((rc = copy_h2d(s, f.r.wc[q], v->wrap[q], ib))) return rc;
    if ((rc = copy_h2d(s, f.r.idx, v->index, ib))) return rc;
    if (v->nbr_off) {
        int maxn = 0;
        for (int i = 0; i < n; ++i) maxn = std::max(maxn, v->nbr_off[i + 1] - v->nbr_off[i]);
        if (maxn > f.nbr_cap) {
            rc = grow_lists(s, maxn, false);
            if (rc) return rc;
            rc = read_header(s, f);
            if (rc) return rc;
        }
        std::vector<int32_t> nn(n), ell((size_t)std::max(maxn, 1) * n, 0);
        for (int i = 0; i < n; ++i) {
            nn[i] = v->nbr_off[i + 1] - v->nbr_off[i];
            for (int q = 0; q < nn[i]; ++q) ell[(size_t)q * n + i] = v->nbr[v->nbr_off[i] + q];
        }
        if ((rc = copy_h2d(s, f.s.nn, nn.data(), ib))) return rc;
        if (maxn > 0)
            HIP_TRY(s, hipMemcpy2D(f.s.nbr, (size_t)f.cap * sizeof(int32_t), ell.data(),
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
    for (int k = 0; k < 3; ++k) d.cell_dim[k] = sc->cell_dim[k];
    d.need_search = 0;
    d.need_remove = 0;
    d.halt = 0;
    d.overflow = 0;
    return write_dyn(s, d);
}

}  // extern "C"
