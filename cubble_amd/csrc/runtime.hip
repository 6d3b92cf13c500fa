// runtime.hip -- the C ABI of include/cubble_amd.h over the engine.
//
// A cubble_sim is an engine holding one foam; a cubble_batch is an engine
// holding many.  The step control flow runs on the device (a one-block
// control kernel per foam decides accept / reject, dt, removal and search),
// so a run of many steps is a queue of identical kernel groups, captured once
// into a HIP graph and replayed, with one host synchronisation per replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "autotune.hpp"
#include "cubble_amd.h"
#include "engine.hpp"
#include "foam.hpp"
#include "host_init.hpp"
#include "launch.hpp"

using namespace cubble_amd;

struct cubble_sim {
    Engine e;
};

struct cubble_batch {
    Engine e;
};

namespace cubble_amd {
// Message of the last failing call without a handle (create, stateless
// calls), returned by cubble_last_error(NULL).
std::string g_create_error;
void set_global_error(const std::string &m) { g_create_error = m; }
}  // namespace cubble_amd

namespace {

const char *kVersion = "cubble_amd 0.2 (gfx950)";

#define R_TRY(eng, call)                                                                \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return (eng).fail(e_ == hipErrorOutOfMemory ? CUBBLE_E_NOMEM : CUBBLE_E_HIP, \
                              std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

uint32_t flags_of(const cubble_integrate_params *ip) {
    uint32_t f = 0;
    if (!ip || ip->use_gas_exchange) f |= kGas;
    if (!ip || ip->increment_path) f |= kPath;
    if (ip && ip->use_flow) f |= kFlow;
    return f;
}

Tuning tuning_of(const cubble_tuning *t) {
    Tuning u;
    if (t) {
        u.ipt = t->bubbles_per_thread;
        if (t->search_tile > 0) u.stile = t->search_tile;
        if (t->graph_steps > 0) u.groups = t->graph_steps;
    }
    return u;
}

int retune(Engine &e, const cubble_tuning *t) {
    if (!t || t->bubbles_per_thread < 0 || t->search_tile < 0 || t->graph_steps < 0 ||
        (t->search_tile > 0 && t->search_tile < kMinSearchTile))
        return e.fail(CUBBLE_E_INVALID, "bad launch shapes");
    return e.retune(tuning_of(t));
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

// Allocate, upload and search the foams; the initial search is not counted.
int build(Engine &e, const cubble_config *cfgs, int n, const Tuning &t) {
    int rc = e.init(0, true);
    if (rc) return rc;
    for (int k = 0; k < n; ++k)
        if ((rc = e.add_foam(cfgs[k], t, nullptr))) return rc;
    if ((rc = e.commit_layout())) return rc;
    if ((rc = e.search_pending())) return rc;
    std::vector<Dyn> d;
    if ((rc = e.read_dyn_all(d))) return rc;
    for (int k = 0; k < n; ++k) {
        d[k].searches = 0;
        if ((rc = e.write_dyn(k, d[k]))) return rc;
    }
    return CUBBLE_OK;
}

int time_kernels(Engine &e, const cubble_integrate_params *ip, int reps, cubble_kernel_times *out) {
    if (!out || reps < 1) return e.fail(CUBBLE_E_INVALID, "bad argument");
    int rc = CUBBLE_OK;
    const uint32_t flags = flags_of(ip);
    const Grid g = e.grid();
    std::vector<Dyn> saved;
    if ((rc = e.read_dyn_all(saved))) return rc;
    launch_set_flags(g, -1, -1, 1);
    // each kernel `reps` times back to back between one pair of events: an
    // event between two kernels costs time of its own (it showed as up to
    // 20% of a pair launch against the profiler's dispatch timestamps).  The
    // attempt kernels recompute the same outputs from the same inputs, so
    // the repetitions do the same work as the first launch.
    hipEvent_t x[5];
    for (auto &v : x) R_TRY(e, hipEventCreate(&v));
    launch_predict(g, flags);  // the inputs of the pair kernel
    R_TRY(e, hipEventRecord(x[0], e.stream));
    for (int r = 0; r < reps; ++r) launch_predict(g, flags);
    R_TRY(e, hipEventRecord(x[1], e.stream));
    // (mixed: the pair kernel alone, without the threshold re-decision pass)
    for (int r = 0; r < reps; ++r) {
        if (CUBBLE_PREC_MIXED == e.precision) launch_pair_mixed(g, flags, false);
        else launch_pair(g, flags);
    }
    R_TRY(e, hipEventRecord(x[2], e.stream));
    // the corrector alone (its inputs: the sums, reduced once before)
    if (CUBBLE_PREC_MIXED == e.precision) launch_pair_mixed(g, flags);
    launch_sums(g);
    R_TRY(e, hipEventRecord(x[3], e.stream));
    for (int r = 0; r < reps; ++r) launch_correct(g, flags, CUBBLE_PREC_MIXED == e.precision, false);
    R_TRY(e, hipEventRecord(x[4], e.stream));
    R_TRY(e, hipStreamSynchronize(e.stream));
    // intervals [x0, x1) predict, [x1, x2) pair, [x3, x4) corrector
    double t[3] = {0.0, 0.0, 0.0};
    const int from[3] = {0, 1, 3};
    for (int q = 0; q < 3; ++q) {
        float ms = 0.f;
        R_TRY(e, hipEventElapsedTime(&ms, x[from[q]], x[from[q] + 1]));
        t[q] = ms;
    }
    for (auto &v : x) (void)hipEventDestroy(v);
    out->predict_ms = t[0] / reps;
    out->pair_ms = t[1] / reps;
    out->correct_ms = t[2] / reps;
    out->reps = reps;
    // the controller did not run: dt and the counters are unchanged; the
    // restore keeps them exactly as they were
    std::vector<Dyn> now;
    if ((rc = e.read_dyn_all(now))) return rc;
    for (size_t k = 0; k < now.size(); ++k) {
        now[k].dt = saved[k].dt;
        now[k].attempts = saved[k].attempts;
        if ((rc = e.write_dyn((int)k, now[k]))) return rc;
    }
    return CUBBLE_OK;
}

// The search phase timed from a copy of the device state, which is put back
// afterwards (a search reorders the bubbles and resets the saved positions).
// Each figure is the average over `reps` back-to-back launches between one
// pair of events on the engine's stream.
int time_search(Engine &e, int reps, cubble_search_times *out) {
    if (!out || reps < 1) return e.fail(CUBBLE_E_INVALID, "bad argument");
    int rc = CUBBLE_OK;
    Snapshot snap;
    if ((rc = e.snapshot(snap))) return rc;
    const Grid g = e.grid();
    std::vector<Dyn> d;
    if ((rc = e.read_dyn_all(d))) {
        e.free_snapshot(snap);
        return rc;
    }
    double bubbles = 0.0;
    for (const Dyn &x : d) bubbles += x.count;
    hipEvent_t ev[2];
    for (auto &v : ev) R_TRY(e, hipEventCreate(&v));
    auto timed = [&](auto &&body, double &ms_out) -> int {
        R_TRY(e, hipEventRecord(ev[0], e.stream));
        for (int r = 0; r < reps; ++r) body();
        R_TRY(e, hipEventRecord(ev[1], e.stream));
        R_TRY(e, hipStreamSynchronize(e.stream));
        float ms = 0.f;
        R_TRY(e, hipEventElapsedTime(&ms, ev[0], ev[1]));
        ms_out = ms / reps;
        return CUBBLE_OK;
    };
    // a whole search of every foam: the flags are set between the searches
    // (k_s_fin clears them); that tiny kernel is timed alone and taken off
    double flags_ms = 0.0, all_ms = 0.0, lists_ms = 0.0, kernel_ms = 0.0, reorder_ms = 0.0,
           idle_ms = 0.0;
    rc = timed([&] { launch_set_flags(g, 1, 0, 1); }, flags_ms);
    if (!rc) rc = timed([&] {
        launch_set_flags(g, 1, 0, 1);
        launch_search(g);
    }, all_ms);
    // the parts: flags set once, the parts do not clear them
    if (!rc) {
        launch_set_flags(g, 1, 0, 1);
        rc = timed([&] { launch_search_part(g, kPartLists); }, lists_ms);
    }
    if (!rc) rc = timed([&] { launch_search_part(g, kPartListKernel); }, kernel_ms);
    if (!rc) rc = timed([&] { launch_search_part(g, kPartReorder); }, reorder_ms);
    // an attempt group's removal and search kernels when no foam needs them
    if (!rc) {
        launch_set_flags(g, 0, 0, 1);
        rc = timed([&] {
            launch_remove(g);
            launch_search(g);
        }, idle_ms);
    }
    for (auto &v : ev) (void)hipEventDestroy(v);
    const int rr = e.restore(snap);
    e.free_snapshot(snap);
    if (rc) return rc;
    if (rr) return rr;
    std::vector<Dyn> after;
    if ((rc = e.read_dyn_all(after))) return rc;
    double pairs = 0.0;
    for (const Dyn &x : after) pairs += x.num_pairs;
    out->search_ms = std::max(0.0, all_ms - flags_ms);
    out->lists_ms = lists_ms;
    out->list_kernel_ms = kernel_ms;
    out->reorder_ms = reorder_ms;
    out->idle_ms = idle_ms;
    out->bubbles = bubbles;
    out->pairs = pairs;
    out->reps = reps;
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

int cubble_create_tuned(const cubble_config *cfg, const cubble_tuning *tuning, cubble_sim **out) {
    if (!out) return CUBBLE_E_INVALID;
    *out = nullptr;
    if (!cfg) {
        g_create_error = "null configuration";
        return CUBBLE_E_INVALID;
    }
    cubble_sim *s = new cubble_sim();
    const int rc = build(s->e, cfg, 1, tuning_of(tuning));
    if (rc) {
        g_create_error = s->e.err;
        delete s;
        return rc;
    }
    *out = s;
    return CUBBLE_OK;
}

int cubble_autotune(cubble_sim *sim, const cubble_integrate_params *ip,
                    const cubble_autotune_opts *opts, cubble_autotune_report *report) {
    if (!sim) return CUBBLE_E_INVALID;
    return autotune(sim->e, flags_of(ip) | kCommit, opts, report);
}

int cubble_retune(cubble_sim *sim, const cubble_tuning *tuning) {
    return sim ? retune(sim->e, tuning) : CUBBLE_E_INVALID;
}

int cubble_create(const cubble_config *cfg, cubble_sim **out) {
    return cubble_create_tuned(cfg, nullptr, out);
}

int cubble_create_from_json(const char *json_path, cubble_sim **out) {
    if (!out) return CUBBLE_E_INVALID;
    *out = nullptr;
    cubble_config cfg;
    std::string err;
    const int rc = parse_config_json(json_path, cfg, err);
    if (rc) {
        g_create_error = err;
        return rc;
    }
    return cubble_create_tuned(&cfg, nullptr, out);
}

void cubble_destroy(cubble_sim *s) { delete s; }

const char *cubble_last_error(const cubble_sim *s) {
    return s ? s->e.err.c_str() : g_create_error.c_str();
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
        for (int q = 0; q < 3; ++q) {
            sc->wall[q] = io.c.wall[q];
            sc->lbb[q] = io.c.lbb[q];
            sc->tfr[q] = io.c.tfr[q];
            sc->interval[q] = io.c.interval[q];
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
    Engine &e = s->e;
    int rc = CUBBLE_OK;
    launch_set_flags(e.grid(), -1, -1, 1);
    e.enqueue_attempt(flags_of(ip));
    R_TRY(e, hipGetLastError());
    std::vector<Dyn> d;
    if ((rc = e.read_dyn_all(d))) return rc;
    fill_result(d[0], res);
    return CUBBLE_OK;
}

int cubble_integrate(cubble_sim *s, const cubble_integrate_params *ip, cubble_step_result *res) {
    if (!s) return CUBBLE_E_INVALID;
    int rc = CUBBLE_OK;
    std::vector<Dyn> d;
    if ((rc = s->e.integrate(flags_of(ip) | kCommit, d))) return rc;
    fill_result(d[0], res);
    return CUBBLE_OK;
}

int cubble_run_steps(cubble_sim *s, const cubble_integrate_params *ip, int64_t num_steps,
                     cubble_step_result *res) {
    if (!s || num_steps < 0) return CUBBLE_E_INVALID;
    int rc = CUBBLE_OK;
    std::vector<Dyn> d;
    if ((rc = s->e.run_steps(flags_of(ip) | kCommit, num_steps, d))) return rc;
    fill_result(d[0], res);
    return CUBBLE_OK;
}

int cubble_search_neighbors(cubble_sim *s) {
    if (!s) return CUBBLE_E_INVALID;
    launch_set_flags(s->e.grid(), 1, -1, 1);
    return s->e.search_pending();
}

int cubble_remove_bubbles(cubble_sim *s) {
    if (!s) return CUBBLE_E_INVALID;
    Engine &e = s->e;
    const Grid g = e.grid();
    launch_set_flags(g, -1, 1, 0);
    launch_remove(g);
    R_TRY(e, hipGetLastError());
    return e.sync();
}

int cubble_get_scalars(cubble_sim *s, cubble_scalars *o) {
    return s ? get_scalars(s->e, 0, o) : CUBBLE_E_INVALID;
}

int cubble_get_state(cubble_sim *s, cubble_state_view *v) {
    return s ? get_state(s->e, 0, v) : CUBBLE_E_INVALID;
}

int cubble_get_attempt(cubble_sim *s, cubble_attempt_view *v) {
    return s ? get_attempt(s->e, 0, v) : CUBBLE_E_INVALID;
}

int cubble_set_state(cubble_sim *s, const cubble_state_view *v, const cubble_scalars *sc) {
    return s ? set_state(s->e, 0, v, sc) : CUBBLE_E_INVALID;
}

int cubble_observe(cubble_sim *s, cubble_observables *out, double *energy) {
    if (!s || !out) return CUBBLE_E_INVALID;
    Engine &e = s->e;
    Foam h;
    int rc = e.read_header(0, h);
    if (rc) return rc;
    const int n = h.d.count;
    // 8 sums, then the per-bubble energies when asked for
    const size_t bytes = sizeof(double) * (kPartD + (energy ? (size_t)n : 0));
    double *buf = static_cast<double *>(e.arena.alloc(bytes));
    if (!buf) return e.fail(CUBBLE_E_NOMEM, "observe: no device memory");
    launch_observe(h, e.dfoams(), buf, energy ? buf + kPartD : nullptr, e.stream);
    double sums[kPartD];
    hipError_t err = hipGetLastError();
    if (hipSuccess == err) err = hipMemcpyAsync(sums, buf, sizeof(sums), hipMemcpyDeviceToHost, e.stream);
    if (hipSuccess == err && energy && n > 0)
        err = hipMemcpyAsync(energy, buf + kPartD, sizeof(double) * n, hipMemcpyDeviceToHost, e.stream);
    if (hipSuccess == err) err = hipStreamSynchronize(e.stream);
    e.arena.release(buf);
    R_TRY(e, err);
    out->total_energy = sums[0];
    out->total_volume = sums[1];
    out->box_volume = h.c.interval[0] * h.c.interval[1] * (3 == h.c.dim ? h.c.interval[2] : 1.0);
    // getAvg divides the sum by the count (src/simulator.cu:1049)
    out->mean_r = n > 0 ? sums[2] / n : 0.0;
    out->mean_path = n > 0 ? sums[3] / n : 0.0;
    for (int q = 0; q < 4; ++q) out->mean_v[q] = n > 0 ? sums[4 + q] / n : 0.0;
    return CUBBLE_OK;
}

int cubble_set_stop_rules(cubble_sim *s, const cubble_stop_rules *r) {
    if (!s || !r) return CUBBLE_E_INVALID;
    if (r->min_count < 0 || r->max_radius < 0 || r->time_frac < 0 || r->time_frac >= 1)
        return s->e.fail(CUBBLE_E_INVALID, "stop rules: negative limit or time fraction outside [0, 1)");
    Foam h;
    int rc = s->e.read_header(0, h);
    if (rc) return rc;
    h.d.count_stop = r->min_count;
    h.d.r_stop = r->max_radius;
    h.d.time_stop = r->use_time ? 1 : 0;
    h.d.stop_int = r->time_int;
    h.d.stop_frac = r->time_frac;
    return s->e.write_dyn(0, h.d);
}

int cubble_reset_print_window(cubble_sim *s) {
    if (!s) return CUBBLE_E_INVALID;
    Engine &e = s->e;
    Foam h;
    int rc = e.read_header(0, h);
    if (rc) return rc;
    const size_t elem = CUBBLE_PREC_MIXED == h.precision ? sizeof(float) : sizeof(double);
    // the error role in the header is current (the search swaps it)
    R_TRY(e, hipMemsetAsync(h.r.err, 0, elem * (size_t)h.cap, e.stream));
    h.d.ts_min = kTsMinEmpty;
    h.d.ts_max = kTsMaxEmpty;
    h.d.ts_sum = 0.0;
    return e.write_dyn(0, h.d);
}

int cubble_rescale_box(cubble_sim *s, double phi, const double rel_dim[3]) {
    if (!s || !rel_dim || !(phi > 0.0)) return CUBBLE_E_INVALID;
    Engine &e = s->e;
    cubble_observables o;
    int rc = cubble_observe(s, &o, nullptr);
    if (rc) return rc;
    if (!(o.total_volume > 0.0)) return e.fail(CUBBLE_E_STATE, "rescale_box: no bubble volume");
    Foam h;
    if ((rc = e.read_header(0, h))) return rc;
    const int D = h.c.dim;
    // reference src/simulator.cu:800-826
    double t = o.total_volume / (phi * rel_dim[0] * rel_dim[1]);
    if (3 == D) {
        t /= rel_dim[2];
        t = std::cbrt(t);
    } else {
        t = std::sqrt(t);
    }
    Consts c = h.c;
    for (int k = 0; k < 3; ++k) {
        c.tfr[k] = k < D ? t * rel_dim[k] : 0.0;
        c.interval[k] = c.tfr[k] - c.lbb[k];
        // the flow box keeps its place relative to the box
        if (k < D) {
            c.flow_lbb[k] = c.interval[k] * ((h.c.flow_lbb[k] - h.c.lbb[k]) / h.c.interval[k]) + c.lbb[k];
            c.flow_tfr[k] = c.interval[k] * ((h.c.flow_tfr[k] - h.c.lbb[k]) / h.c.interval[k]) + c.lbb[k];
        }
    }
    double vbox = c.interval[0] * c.interval[1];
    if (3 == D) vbox *= c.interval[2];
    const double mult = phi * vbox / kPi;
    c.vol_mul = 3 == D ? std::cbrt(0.75 * mult) : std::sqrt(mult);
    launch_rescale(h, e.dfoams(), c, e.stream);
    R_TRY(e, hipGetLastError());
    R_TRY(e, hipStreamSynchronize(e.stream));
    R_TRY(e, hipMemcpy(&e.dfoams()[0].c, &c, sizeof(Consts), hipMemcpyHostToDevice));
    e.foams[0].h.c = c;
    launch_set_flags(e.grid(), 1, -1, 1);
    return e.search_pending();
}

int cubble_bootstrap_velocities(cubble_sim *s) {
    if (!s) return CUBBLE_E_INVALID;
    Engine &e = s->e;
    Foam h;
    int rc = e.read_header(0, h);
    if (rc) return rc;
    if (h.d.attempts != 0 || h.d.steps != 0)
        return e.fail(CUBBLE_E_STATE, "bootstrap_velocities: only before the first attempt");
    launch_bootstrap(e.grid(), CUBBLE_PREC_MIXED == e.precision);
    R_TRY(e, hipGetLastError());
    return e.sync();
}

int cubble_reset_clock(cubble_sim *s, int32_t clear_wraps) {
    if (!s) return CUBBLE_E_INVALID;
    Engine &e = s->e;
    Foam h;
    int rc = e.read_header(0, h);
    if (rc) return rc;
    if (clear_wraps)
        for (int q = 0; q < h.c.dim; ++q)
            R_TRY(e, hipMemsetAsync(h.r.wc[q], 0, sizeof(int32_t) * (size_t)h.cap, e.stream));
    h.d.time_int = 0;
    h.d.time_frac = 0.0;
    return e.write_dyn(0, h.d);
}

int cubble_time_kernels(cubble_sim *s, const cubble_integrate_params *ip, int32_t reps,
                        cubble_kernel_times *out) {
    return s ? time_kernels(s->e, ip, reps, out) : CUBBLE_E_INVALID;
}

int cubble_time_search(cubble_sim *s, int32_t reps, cubble_search_times *out) {
    return s ? time_search(s->e, reps, out) : CUBBLE_E_INVALID;
}

// ---------------------------------------------------------------- batches
int cubble_batch_create(const cubble_config *cfgs, int32_t n, const cubble_tuning *tuning,
                        cubble_batch **out) {
    if (!out) return CUBBLE_E_INVALID;
    *out = nullptr;
    if (!cfgs || n < 1) {
        g_create_error = "a batch needs at least one configuration";
        return CUBBLE_E_INVALID;
    }
    cubble_batch *b = new cubble_batch();
    const int rc = build(b->e, cfgs, n, tuning_of(tuning));
    if (rc) {
        g_create_error = b->e.err;
        delete b;
        return rc;
    }
    *out = b;
    return CUBBLE_OK;
}

int cubble_batch_size(const cubble_batch *b) { return b ? (int)b->e.foams.size() : 0; }

int cubble_batch_integrate(cubble_batch *b, const cubble_integrate_params *ip,
                           cubble_step_result *per_foam) {
    if (!b) return CUBBLE_E_INVALID;
    int rc = CUBBLE_OK;
    std::vector<Dyn> d;
    if ((rc = b->e.integrate(flags_of(ip) | kCommit, d))) return rc;
    if (per_foam)
        for (size_t k = 0; k < d.size(); ++k) fill_result(d[k], per_foam + k);
    return CUBBLE_OK;
}

int cubble_batch_run_steps(cubble_batch *b, const cubble_integrate_params *ip, int64_t num_steps,
                           cubble_step_result *per_foam) {
    if (!b || num_steps < 0) return CUBBLE_E_INVALID;
    int rc = CUBBLE_OK;
    std::vector<Dyn> d;
    if ((rc = b->e.run_steps(flags_of(ip) | kCommit, num_steps, d))) return rc;
    if (per_foam)
        for (size_t k = 0; k < d.size(); ++k) fill_result(d[k], per_foam + k);
    return CUBBLE_OK;
}

int cubble_batch_get_scalars(cubble_batch *b, int32_t k, cubble_scalars *out) {
    return b ? get_scalars(b->e, k, out) : CUBBLE_E_INVALID;
}

int cubble_batch_get_state(cubble_batch *b, int32_t k, cubble_state_view *view) {
    return b ? get_state(b->e, k, view) : CUBBLE_E_INVALID;
}

int cubble_batch_time_kernels(cubble_batch *b, const cubble_integrate_params *ip, int32_t reps,
                              cubble_kernel_times *out) {
    return b ? time_kernels(b->e, ip, reps, out) : CUBBLE_E_INVALID;
}

int cubble_batch_time_search(cubble_batch *b, int32_t reps, cubble_search_times *out) {
    return b ? time_search(b->e, reps, out) : CUBBLE_E_INVALID;
}

int cubble_batch_autotune(cubble_batch *b, const cubble_integrate_params *ip,
                          const cubble_autotune_opts *opts, cubble_autotune_report *report) {
    if (!b) return CUBBLE_E_INVALID;
    return autotune(b->e, flags_of(ip) | kCommit, opts, report);
}

int cubble_batch_retune(cubble_batch *b, const cubble_tuning *tuning) {
    return b ? retune(b->e, tuning) : CUBBLE_E_INVALID;
}

const char *cubble_batch_last_error(const cubble_batch *b) {
    return b ? b->e.err.c_str() : g_create_error.c_str();
}

void cubble_batch_destroy(cubble_batch *b) { delete b; }

}  // extern "C"
