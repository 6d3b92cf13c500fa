// server.hip -- request loop for foam simulations with continuous batching
// (the serving C ABI of include/cubble_amd.h).
//
// Requests queue in arrival order.  At every round boundary the loop admits
// queued requests whose arrival time has passed, head of line first, while
// their device footprint fits the fixed-size device arena (the HBM budget)
// and a resident slot is free; admitted foams are drawn on the host,
// uploaded, searched and stay resident.  A round advances every resident
// foam by up to `round_steps` accepted steps in the batched, graph-replayed
// engine; each foam stops on the device at its own step budget or when its
// bubble count falls to the dissolution threshold.  Finished foams are
// retired (their memory returns to the arena) before the next admission.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
#include <vector>

#include "cubble_amd.h"
#include "engine.hpp"
#include "host_init.hpp"

using namespace cubble_amd;

namespace {

struct KeptState {
    cubble_scalars sc{};
    std::vector<double> x, v, vo, saved, path, err;
    std::vector<int32_t> wc, idx, nbr_off, nbr, deg;
};

struct Request {
    cubble_request_result r{};
    cubble_config cfg{};
    int64_t max_steps = 0;
    double dissolve = 0.0;
    int32_t count_stop = 0;
    bool drawn = false;
    std::vector<double> x4;  // host draw, kept until admission
    InitOut io{};
    KeptState kept;
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct cubble_server {
    Engine e;
    cubble_server_opts opts{};
    std::vector<Request> reqs;
    std::deque<int64_t> queue;      // request ids in arrival order
    std::vector<int64_t> resident;  // engine foam index -> request id
};

namespace {

int keep_state(cubble_server *s, int k, Request &q) {
    KeptState &ks = q.kept;
    int rc = get_scalars(s->e, k, &ks.sc);
    if (rc) return rc;
    const size_t n = (size_t)ks.sc.count;
    ks.x.assign(4 * n, 0.0);
    ks.v.assign(4 * n, 0.0);
    ks.vo.assign(4 * n, 0.0);
    ks.saved.assign(4 * n, 0.0);
    ks.path.assign(n, 0.0);
    ks.err.assign(n, 0.0);
    ks.wc.assign(3 * n, 0);
    ks.idx.assign(n, 0);
    ks.deg.assign(n, 0);
    ks.nbr_off.assign(n + 1, 0);
    ks.nbr.assign((size_t)ks.sc.nbr_cap * n + 1, 0);
    cubble_state_view v{};
    for (int c = 0; c < 4; ++c) {
        v.x[c] = ks.x.data() + c * n;
        v.v[c] = ks.v.data() + c * n;
        v.vo[c] = ks.vo.data() + c * n;
        v.saved[c] = ks.saved.data() + c * n;
    }
    for (int c = 0; c < 3; ++c) v.wrap[c] = ks.wc.data() + c * n;
    v.path = ks.path.data();
    v.error = ks.err.data();
    v.index = ks.idx.data();
    v.degree = ks.deg.data();
    v.nbr_off = ks.nbr_off.data();
    v.nbr = ks.nbr.data();
    v.nbr_cap = (int64_t)ks.nbr.size();
    rc = get_state(s->e, k, &v);
    if (rc) return rc;
    ks.nbr.resize((size_t)ks.nbr_off[n]);
    return CUBBLE_OK;
}

// Admit queued requests that have arrived, in order, while they fit.
int admit(cubble_server *s, double t, bool &changed) {
    Engine &e = s->e;
    const int max_active = s->opts.max_active;
    Tuning tune;
    while (!s->queue.empty()) {
        Request &q = s->reqs[s->queue.front()];
        if (q.r.arrival_s > t) break;
        if (max_active > 0 && (int)s->resident.size() >= max_active) break;
        if (!q.drawn) {
            std::string err;
            const int rc = generate_foam(q.cfg, q.x4, q.io, err);
            if (rc) {  // a malformed request fails alone
                q.r.status = rc;
                q.r.finish_s = t;
                s->queue.pop_front();
                continue;
            }
            q.drawn = true;
        }
        const size_t need = foam_footprint(q.cfg.dim, q.cfg.precision, q.cfg.num_start,
                                           q.cfg.nbr_cap, q.io.cell_cap, tune);
        // head-of-line: wait for memory; a request that can never fit fails
        if (need > e.arena.capacity()) {
            q.r.status = CUBBLE_E_NOMEM;
            q.r.finish_s = t;
            s->queue.pop_front();
            continue;
        }
        if (need + need / 8 > e.arena.largest_free()) break;
        int k = -1;
        int rc = e.add_generated(q.cfg, tune, q.x4, q.io, &k);
        if (CUBBLE_E_NOMEM == rc) break;
        if (rc) return rc;
        Foam &h = e.foams[k].h;
        h.d.count_stop = q.count_stop;
        q.x4.clear();
        q.x4.shrink_to_fit();
        q.r.admit_s = t;
        q.r.status = 1;
        s->resident.push_back(q.r.id);
        s->queue.pop_front();
        changed = true;
    }
    return CUBBLE_OK;
}

}  // namespace

extern "C" {

int64_t cubble_foam_footprint(const cubble_config *cfg) {
    if (!cfg) return -1;
    std::vector<double> x4;
    InitOut io;
    std::string err;
    if (generate_foam(*cfg, x4, io, err)) return -1;
    return (int64_t)foam_footprint(cfg->dim, cfg->precision, cfg->num_start, cfg->nbr_cap,
                                   io.cell_cap, Tuning());
}

int cubble_server_create(const cubble_server_opts *opts, cubble_server **out) {
    if (!out || !opts || opts->hbm_budget_bytes == 0) return CUBBLE_E_INVALID;
    *out = nullptr;
    cubble_server *s = new cubble_server();
    s->opts = *opts;
    if (s->opts.round_steps <= 0) s->opts.round_steps = 50;
    s->e.precision = opts->precision;
    s->e.isolate_failures = true;
    const int rc = s->e.init((size_t)opts->hbm_budget_bytes, false);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return CUBBLE_OK;
}

int cubble_server_submit(cubble_server *s, const cubble_config *cfg, int64_t max_steps,
                         double dissolve_fraction, double arrival_s, int64_t *id) {
    if (!s || !cfg || max_steps <= 0) return CUBBLE_E_INVALID;
    if (cfg->precision != s->opts.precision)
        return s->e.fail(CUBBLE_E_INVALID, "request precision differs from the server's");
    // every resident foam runs under one set of launch flags, without flow
    if (cfg->flow)
        return s->e.fail(CUBBLE_E_UNSUPPORTED, "the serving loop does not run the imposed flow");
    Request q;
    q.cfg = *cfg;
    q.max_steps = max_steps;
    q.dissolve = dissolve_fraction;
    q.count_stop = dissolve_fraction > 0.0
                       ? (int32_t)std::floor((1.0 - dissolve_fraction) * cfg->num_start)
                       : 0;
    q.r.id = (int64_t)s->reqs.size();
    q.r.arrival_s = arrival_s;
    q.r.admit_s = -1.0;
    q.r.finish_s = -1.0;
    q.r.start_count = cfg->num_start;
    // keep the queue ordered by arrival
    s->reqs.push_back(q);
    auto it = s->queue.end();
    while (it != s->queue.begin() && s->reqs[*(it - 1)].r.arrival_s > arrival_s) --it;
    s->queue.insert(it, q.r.id);
    if (id) *id = q.r.id;
    return CUBBLE_OK;
}

int cubble_server_run(cubble_server *s, cubble_server_stats *stats) {
    if (!s) return CUBBLE_E_INVALID;
    Engine &e = s->e;
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t flags = kGas | kPath | kCommit;
    int64_t rounds = 0;
    int max_resident = 0;
    double peak = 0.0;
    int rc = CUBBLE_OK;
    while (!s->queue.empty() || !s->resident.empty()) {
        bool changed = false;
        if ((rc = admit(s, now_s(t0), changed))) return rc;
        if (changed) {
            if ((rc = e.commit_layout())) return rc;
            // first search of the new foams only (their flag is set)
            std::vector<Dyn> before;
            if ((rc = e.read_dyn_all(before))) return rc;
            if ((rc = e.search_pending())) return rc;
            std::vector<Dyn> after;
            if ((rc = e.read_dyn_all(after))) return rc;
            for (size_t k = 0; k < after.size(); ++k)
                if (after[k].searches != before[k].searches && before[k].steps == 0) {
                    after[k].searches = 0;
                    if ((rc = e.write_dyn((int)k, after[k]))) return rc;
                }
        }
        max_resident = std::max(max_resident, (int)s->resident.size());
        peak = std::max(peak, (double)e.arena.in_use());
        if (s->resident.empty()) {
            // admission may have failed the last queued request
            if (s->queue.empty()) break;
            // idle until the next arrival
            const double wait = s->reqs[s->queue.front()].r.arrival_s - now_s(t0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            continue;
        }
        std::vector<Dyn> d;
        if ((rc = e.read_dyn_all(d))) return rc;
        std::vector<int64_t> budget(s->resident.size());
        for (size_t k = 0; k < s->resident.size(); ++k) {
            const Request &q = s->reqs[s->resident[k]];
            budget[k] = std::min<int64_t>(s->opts.round_steps, q.max_steps - d[k].steps);
        }
        if ((rc = e.run_steps(flags, s->opts.round_steps, d, &budget))) return rc;
        ++rounds;
        // retire finished foams, highest index first
        bool dropped = false;
        const double t = now_s(t0);
        for (int k = (int)s->resident.size() - 1; k >= 0; --k) {
            Request &q = s->reqs[s->resident[k]];
            if (e.foams[k].failed) {
                // its lists could not grow in the budget: it fails alone
                q.r.status = e.foams[k].failed;
                q.r.finish_s = t;
                q.r.final_count = d[k].count;
                q.r.steps = d[k].steps;
                q.r.bubble_steps = d[k].bubble_steps;
                if ((rc = e.drop_foam(k))) return rc;
                s->resident.erase(s->resident.begin() + k);
                dropped = true;
                continue;
            }
            const bool done = d[k].steps >= q.max_steps ||
                              (q.count_stop > 0 && d[k].count <= q.count_stop);
            if (!done) continue;
            q.r.status = 2;
            q.r.finish_s = t;
            q.r.final_count = d[k].count;
            q.r.steps = d[k].steps;
            q.r.bubble_steps = d[k].bubble_steps;
            if (s->opts.keep_states && (rc = keep_state(s, k, q))) return rc;
            if ((rc = e.drop_foam(k))) return rc;
            s->resident.erase(s->resident.begin() + k);
            dropped = true;
        }
        if (dropped && (rc = e.commit_layout())) return rc;
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        std::vector<double> lat;
        double bs = 0.0;
        for (const Request &q : s->reqs) {
            if (q.r.status != 2) continue;
            lat.push_back(q.r.finish_s - q.r.arrival_s);
            bs += q.r.bubble_steps;
        }
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
            if (lat.empty()) return 0.0;
            const size_t i = std::min(lat.size() - 1, (size_t)std::ceil(p * lat.size()) - 1);
            return lat[i];
        };
        stats->requests = (int64_t)lat.size();
        stats->rounds = rounds;
        stats->max_resident = max_resident;
        stats->wall_s = now_s(t0);
        stats->bubble_steps = bs;
        stats->throughput = bs / std::max(stats->wall_s, 1e-9);
        stats->p50_latency_s = pct(0.50);
        stats->p99_latency_s = pct(0.99);
        stats->peak_bytes = peak;
    }
    return CUBBLE_OK;
}

int cubble_server_result(cubble_server *s, int64_t id, cubble_request_result *out) {
    if (!s || !out || id < 0 || id >= (int64_t)s->reqs.size()) return CUBBLE_E_INVALID;
    *out = s->reqs[id].r;
    return CUBBLE_OK;
}

int cubble_server_get_state(cubble_server *s, int64_t id, cubble_state_view *v,
                            cubble_scalars *sc) {
    if (!s || id < 0 || id >= (int64_t)s->reqs.size()) return CUBBLE_E_INVALID;
    const Request &q = s->reqs[id];
    if (q.r.status != 2 || q.kept.x.empty())
        return s->e.fail(CUBBLE_E_STATE, "no kept state for this request");
    const KeptState &ks = q.kept;
    const size_t n = (size_t)ks.sc.count;
    if (sc) *sc = ks.sc;
    if (!v) return CUBBLE_OK;
    auto cp = [](void *dst, const void *src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    for (int c = 0; c < 4; ++c) {
        cp(v->x[c], ks.x.data() + c * n, n * sizeof(double));
        cp(v->v[c], ks.v.data() + c * n, n * sizeof(double));
        cp(v->vo[c], ks.vo.data() + c * n, n * sizeof(double));
        cp(v->saved[c], ks.saved.data() + c * n, n * sizeof(double));
    }
    for (int c = 0; c < 3; ++c) cp(v->wrap[c], ks.wc.data() + c * n, n * sizeof(int32_t));
    cp(v->path, ks.path.data(), n * sizeof(double));
    cp(v->error, ks.err.data(), n * sizeof(double));
    cp(v->index, ks.idx.data(), n * sizeof(int32_t));
    cp(v->degree, ks.deg.data(), n * sizeof(int32_t));
    cp(v->nbr_off, ks.nbr_off.data(), (n + 1) * sizeof(int32_t));
    if (v->nbr) {
        if ((int64_t)ks.nbr.size() > v->nbr_cap)
            return s->e.fail(CUBBLE_E_CAPACITY, "nbr buffer too small");
        cp(v->nbr, ks.nbr.data(), ks.nbr.size() * sizeof(int32_t));
    }
    return CUBBLE_OK;
}

const char *cubble_server_last_error(const cubble_server *s) { return s ? s->e.err.c_str() : ""; }

void cubble_server_destroy(cubble_server *s) { delete s; }

}  // extern "C"
