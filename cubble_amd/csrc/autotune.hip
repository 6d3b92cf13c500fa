// autotune.hip -- picks the launch shapes of a set of foams by timing them.
//
// Three shapes matter on this path: bubbles per thread of the bubble-space
// kernels (block count against per-thread work and the number of block
// partials to reduce), bubbles per neighbour-search workgroup (the search
// tile: small tiles give the few searching foams of a batch enough
// workgroups, large ones cheapen the idle launch over every foam's tiles),
// and attempt groups per
// captured graph replay (host synchronisations against the groups a replay
// wastes after every foam reached its step limit).
//
// The tuner copies the whole device state aside, then for each candidate
// restores it, applies the shapes, primes the captured graph, restores again
// and times `trial_steps` accepted steps, so every candidate does the same
// work from the same state.  Coordinate descent over the three shapes, each
// candidate timed `repeats` times (fastest kept).  The state is restored at
// the end: tuning does not advance the simulation.  Results go to an
// optional JSON cache keyed by the device and the make-up of the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "autotune.hpp"
#include "host_init.hpp"

namespace cubble_amd {

namespace {

int log2_bucket(double v) { return v <= 1.0 ? 0 : (int)std::floor(std::log2(v)); }

std::string cache_key(const Engine &e) {
    hipDeviceProp_t prop{};
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDeviceProperties(&prop, e.device) == hipSuccess) {
        arch = prop.gcnArchName;
        const size_t colon = arch.find(':');
        if (colon != std::string::npos) arch = arch.substr(0, colon);
        cus = prop.multiProcessorCount;
    }
    int n2 = 0, n3 = 0;
    double total = 0.0, largest = 0.0;
    for (const FoamRec &r : e.foams) {
        (3 == r.h.c.dim ? n3 : n2) += 1;
        total += r.h.cap;
        largest = std::max(largest, (double)r.h.cap);
    }
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s/%dcu/%s/2d%d/3d%d/total2^%d/max2^%d", arch.c_str(), cus,
                  CUBBLE_PREC_MIXED == e.precision ? "mixed" : "fp64", log2_bucket(n2 + 0.5) ,
                  log2_bucket(n3 + 0.5), log2_bucket(total), log2_bucket(largest));
    return buf;
}

// Accepted bubble-steps per second of `steps` steps from the snapshot.
int time_shapes(Engine &e, const Snapshot &snap, uint32_t flags, const Tuning &t, int steps,
                double &rate) {
    int rc;
    std::vector<Dyn> d0, d1;
    if ((rc = e.restore(snap)) || (rc = e.retune(t))) return rc;
    // prime: capture and replay the graph once
    if ((rc = e.run_steps(flags, std::max(1, t.groups), d1))) return rc;
    if ((rc = e.restore(snap)) || (rc = e.read_dyn_all(d0))) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = e.run_steps(flags, steps, d1))) return rc;
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double work = 0.0;
    for (size_t k = 0; k < d1.size(); ++k) work += (double)(d1[k].bubble_steps - d0[k].bubble_steps);
    rate = s > 0.0 ? work / s : 0.0;
    return CUBBLE_OK;
}

}  // namespace

int autotune(Engine &e, uint32_t flags, const cubble_autotune_opts *o, cubble_autotune_report *rep) {
    if (e.foams.empty()) return e.fail(CUBBLE_E_INVALID, "nothing to tune");
    const int steps = o && o->trial_steps > 0 ? o->trial_steps : 64;
    const int repeats = o && o->repeats > 0 ? o->repeats : 2;
    const char *cache = o ? o->cache_path : nullptr;
    const std::string key = cache_key(e);
    cubble_autotune_report r{};
    int shape[3];
    if (cache && (!o || !o->ignore_cache) && tuning_cache_lookup(cache, key, shape)) {
        Tuning t;
        t.ipt = shape[0];
        t.stile = shape[1];
        t.groups = shape[2];
        const int rc = e.retune(t);
        if (rc) return rc;
        r.chosen.bubbles_per_thread = t.ipt;
        r.chosen.search_tile = t.stile;
        r.chosen.graph_steps = t.groups;
        r.from_cache = 1;
        if (rep) *rep = r;
        return CUBBLE_OK;
    }

    Snapshot snap;
    int rc = e.snapshot(snap);
    if (rc) return rc;
    const Tuning start = e.tuning();
    auto measure = [&](const Tuning &t, double &best) -> int {
        best = 0.0;
        for (int k = 0; k < repeats; ++k) {
            double v = 0.0;
            const int c = time_shapes(e, snap, flags, t, steps, v);
            if (c) return c;
            best = std::max(best, v);
            ++r.candidates_timed;
        }
        return CUBBLE_OK;
    };

    Tuning best = start;
    double best_rate = 0.0;
    if ((rc = measure(best, best_rate))) {
        e.free_snapshot(snap);
        return rc;
    }
    r.start_rate = best_rate;
    const std::vector<int> ipts = {1, 2, 4, 8};
    const std::vector<int> tiles = {32, 64, 128, 256, 512};
    const std::vector<int> groups = {4, 8, 16, 32};
    for (int axis = 0; axis < 3 && !rc; ++axis) {
        const std::vector<int> &cand = 0 == axis ? ipts : 1 == axis ? tiles : groups;
        const Tuning base = best;
        for (int v : cand) {
            Tuning t = base;
            int &field = 0 == axis ? t.ipt : 1 == axis ? t.stile : t.groups;
            const int cur = 0 == axis ? base.ipt : 1 == axis ? base.stile : base.groups;
            if (v == cur) continue;
            double rate = 0.0;
            field = v;
            if ((rc = measure(t, rate))) break;
            if (rate > best_rate) {
                best_rate = rate;
                best = t;
            }
        }
    }
    if (!rc) rc = e.restore(snap);
    if (!rc) rc = e.retune(best);
    e.free_snapshot(snap);
    if (rc) return rc;
    r.chosen.bubbles_per_thread = best.ipt;
    r.chosen.search_tile = best.stile;
    r.chosen.graph_steps = best.groups;
    r.best_rate = best_rate;
    if (cache) {
        shape[0] = best.ipt;
        shape[1] = best.stile;
        shape[2] = best.groups;
        std::string err;
        if ((rc = tuning_cache_store(cache, key, shape, best_rate, err))) return e.fail(rc, err);
    }
    if (rep) *rep = r;
    return CUBBLE_OK;
}

}  // namespace cubble_amd
