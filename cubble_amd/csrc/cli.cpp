// cli.cpp -- drop-in command line entry, cubble_run(argc, argv), and the
// one-line main of cubble-cli.  Mirrors cubble::run (reference
// src/simulator.cu:1161-1186), the velocity bootstrap and stabilisation
// phase of init() (src/simulator.cu:739-791, :848-955) and simulate()
// (src/simulator.cu:958-1157): argument check, device check, stabilisation
// with its table, the integration loop and its stop rule, the printed table,
// results.dat and the snapshot CSVs.  A lattice start (the default of a file
// without an "init" object, as the reference's own input) follows init: the
// lattice in its own box, the velocity bootstrap, five stabilisation rounds,
// the box rescale to the target gas fraction with a new search, five more
// rounds (src/simulator.cu:576-606, :739-846).  A uniform start ("init":
// {"placement": "uniform"}) is drawn at the target gas fraction directly and
// skips the rescale.
//
// The reference's host loop checks after every accepted step whether to
// print, snapshot or stop.  Here the steps run on the device
// (cubble_run_steps) under stop rules set to the next of those events, so
// each run ends at the accepted step where the reference acts, and every row
// and file is taken from the state of that step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "cubble_amd.h"
#include "host_init.hpp"

namespace {

// Most accepted steps per device run between two looks at the state.
constexpr int64_t kChunk = 1000;

// A scheduled time, split as the reference compares it against its
// (integer, fraction) clock (src/simulator.cu:1025-1034).
struct SplitTime {
    uint64_t i;
    double f;
};

SplitTime split(double t) {
    const uint64_t i = (uint64_t)t;
    return {i, t - (double)i};
}

bool reached(const cubble_scalars &s, SplitTime t) {
    return s.time_int > t.i || (s.time_int == t.i && s.time_frac >= t.f);
}

bool earlier(SplitTime a, SplitTime b) { return a.i < b.i || (a.i == b.i && a.f < b.f); }

int fail(cubble_sim *sim) {
    std::fprintf(stderr, "error: %s\n", cubble_last_error(sim));
    cubble_destroy(sim);
    return EXIT_FAILURE;
}

// stabilize() (reference src/simulator.cu:339-360): `steps` accepted steps
// without gas exchange, flow or path, with the energy before and after.
int stabilize(cubble_sim *sim, int64_t steps, double &e1, double &e2) {
    cubble_observables o;
    int rc = cubble_observe(sim, &o, nullptr);
    if (rc) return rc;
    e1 = o.total_energy;
    const cubble_integrate_params ip = {0, 0, 0};
    cubble_step_result res;
    if ((rc = cubble_run_steps(sim, &ip, steps, &res))) return rc;
    if ((rc = cubble_observe(sim, &o, nullptr))) return rc;
    e2 = o.total_energy;
    return CUBBLE_OK;
}

// Five stabilisation rounds (src/simulator.cu:789-791, :848-850).
int five_rounds(cubble_sim *sim, int64_t steps, const char *when) {
    std::printf("Stabilizing a few rounds after %s\n", when);
    double e1 = 0.0, e2 = 0.0;
    for (int k = 0; k < 5; ++k) {
        const int rc = stabilize(sim, steps, e1, e2);
        if (rc) return rc;
    }
    return CUBBLE_OK;
}

// The box rescale of init (src/simulator.cu:793-846) for a foam placed on
// the lattice in its own box.
int scale_box(cubble_sim *sim, const cubble_config &cfg) {
    std::printf("Scaling the simulation box\n");
    cubble_observables o;
    int rc = cubble_observe(sim, &o, nullptr);
    if (rc) return rc;
    std::printf("Current phi: %.9g, target phi: %.9g\n", o.total_volume / o.box_volume, cfg.phi);
    if ((rc = cubble_rescale_box(sim, cfg.phi, cfg.rel_dim)) ||
        (rc = cubble_observe(sim, &o, nullptr)))
        return rc;
    std::printf("Current phi: %.9g, target phi: %.9g\n", o.total_volume / o.box_volume, cfg.phi);
    std::printf("Neighbor search after scaling\n");
    return CUBBLE_OK;
}

// The stabilisation loop of init() (src/simulator.cu:852-899): rounds until
// the energy change per scaled time is below maxDeltaEnergy (or small for a
// low energy), at most about 500, printing the reference's table.
// `searches0` returns the search count the reference's counter starts the
// first integration row from (it is not reset after the last round).
int stabilization_phase(cubble_sim *sim, int64_t steps, double max_de, double time_scale,
                        int64_t &searches0) {
    double e1 = 0.0, e2 = 0.0;
    int rc = CUBBLE_OK;
    std::printf("\n=============\nStabilization\n=============\n");
    std::printf("%-7s %-11s %-11s %-11s %-9s\n", "#steps", "dE", "e1", "e2", "#searches");
    cubble_scalars sc;
    if ((rc = cubble_get_scalars(sim, &sc))) return rc;
    searches0 = sc.searches;
    const int failsafe = 500;
    for (int num = 0;; ++num) {
        if ((rc = cubble_reset_clock(sim, 0)) || (rc = stabilize(sim, steps, e1, e2)) ||
            (rc = cubble_get_scalars(sim, &sc)))
            return rc;
        const double time = ((double)sc.time_int + sc.time_frac) * time_scale;
        double de = std::abs(e2 - e1);
        if (de > 0.0) de *= 2.0 / ((e2 + e1) * time);
        if (de < max_de || (e2 < 1.0 && de < 0.1)) {
            std::printf("Final energies:");
            std::printf("\nbefore: %9.5e", e1);
            std::printf("\nafter: %9.5e", e2);
            std::printf("\ndelta: %9.5e", de);
            std::printf("\ntime: %9.5g\n", time);
            break;
        } else if (num > failsafe) {
            std::printf("Over %d steps taken and required delta energy not reached. "
                        "Constraints might be too strict.\n",
                        num);
            break;
        }
        std::printf("%-7d ", (int)((num + 1) * steps));
        std::printf("%-9.5e ", de);
        std::printf("%-9.5e ", e1);
        std::printf("%-9.5e ", e2);
        std::printf("%-9ld\n", (long)(sc.searches - searches0));
        searches0 = sc.searches;
    }
    return CUBBLE_OK;
}

// Snapshot files <name>.csv.<k> (reference saveSnapshot, src/simulator.cu:387-512).
class Snapshots {
  public:
    explicit Snapshots(std::string name) : name_(std::move(name)) {}

    // Starting positions, by original index (src/simulator.cu:901-936).
    int start(cubble_sim *sim, const cubble_scalars &sc) {
        if (read(sim, sc)) return CUBBLE_E_HIP;
        int32_t top = 0;
        for (int32_t v : idx_) top = v > top ? v : top;
        for (auto *p : {&px_, &py_, &pz_}) p->assign((size_t)top + 1, 0.0);
        for (size_t i = 0; i < idx_.size(); ++i) {
            px_[idx_[i]] = x_[0][i];
            py_[idx_[i]] = x_[1][i];
            pz_[idx_[i]] = x_[2][i];
        }
        return CUBBLE_OK;
    }

    int write(cubble_sim *sim, const cubble_scalars &sc) {
        if (read(sim, sc)) return CUBBLE_E_HIP;
        std::stringstream ss;
        ss << name_ << ".csv." << num_++;
        std::ofstream file(ss.str().c_str(), std::ios::out);
        if (!file.is_open()) return CUBBLE_OK;  // the reference skips the file silently
        const double *L = sc.interval;
        file << "x,y,z,r,vx,vy,vz,vtot,vr,path,energy,displacement,error,index\n";
        for (size_t i = 0; i < idx_.size(); ++i) {
            const int32_t ind = idx_[i];
            const double xi = x_[0][i], yi = x_[1][i], zi = x_[2][i];
            const double vx = v_[0][i], vy = v_[1][i], vz = v_[2][i];
            // displacement since the previous snapshot, one period folded
            // back per axis (src/simulator.cu:455-468)
            double dx = std::abs(xi - px_[ind]);
            dx = dx > 0.5 * L[0] ? dx - L[0] : dx;
            double dy = std::abs(yi - py_[ind]);
            dy = dy > 0.5 * L[1] ? dy - L[1] : dy;
            double dz = std::abs(zi - pz_[ind]);
            dz = dz > 0.5 * L[2] ? dz - L[2] : dz;
            file << xi - 0.5 * L[0] << "," << yi - 0.5 * L[1] << "," << zi - 0.5 * L[2] << ","
                 << x_[3][i] << "," << vx << "," << vy << "," << vz << ","
                 << std::sqrt(vx * vx + vy * vy + vz * vz) << "," << v_[3][i] << "," << path_[i]
                 << "," << energy_[i] << "," << std::sqrt(dx * dx + dy * dy + dz * dz) << ","
                 << err_[i] << "," << ind << "\n";
            px_[ind] = xi;
            py_[ind] = yi;
            pz_[ind] = zi;
        }
        return CUBBLE_OK;
    }

  private:
    int read(cubble_sim *sim, const cubble_scalars &sc) {
        const size_t n = (size_t)sc.count;
        cubble_state_view v = {};
        for (int q = 0; q < 4; ++q) {
            x_[q].assign(n, 0.0);
            v_[q].assign(n, 0.0);
            v.x[q] = x_[q].data();
            v.v[q] = v_[q].data();
        }
        path_.assign(n, 0.0);
        err_.assign(n, 0.0);
        idx_.assign(n, 0);
        energy_.assign(n, 0.0);
        v.path = path_.data();
        v.error = err_.data();
        v.index = idx_.data();
        if (cubble_get_state(sim, &v) != CUBBLE_OK) return 1;
        cubble_observables o;
        return cubble_observe(sim, &o, energy_.data()) != CUBBLE_OK;
    }

    std::string name_;
    int num_ = 0;
    std::vector<double> px_, py_, pz_;
    std::vector<double> x_[4], v_[4], path_, err_, energy_;
    std::vector<int32_t> idx_;
};

}  // namespace

extern "C" int cubble_run(int argc, char **argv) {
    if (argc != 2) {
        std::printf(
            "Usage: %s inputFile\n\twhere inputFile is the name of the (.json) file containing the "
            "simulation input.\n",
            argc > 0 ? argv[0] : "cubble-cli");
        return EXIT_FAILURE;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        std::printf("No HIP capable devices found.\n");
        return EXIT_FAILURE;
    }
    cubble_config cfg;
    if (cubble_config_from_json(argv[1], &cfg) != CUBBLE_OK) return fail(nullptr);
    cubble_amd::RunSettings rs;
    std::string err;
    if (cubble_amd::parse_run_json(argv[1], rs, err) != CUBBLE_OK) {
        std::fprintf(stderr, "error: %s\n", err.c_str());
        return EXIT_FAILURE;
    }
    const double snap_freq = rs.snapshot_frequency;
    // a lattice start is created in the lattice's own box and scaled to the
    // target gas fraction after the first stabilisation rounds, as init does
    const bool lattice = 1 == cfg.placement;
    cubble_config created = cfg;
    if (lattice) created.placement = 2;
    cubble_sim *sim = nullptr;
    if (cubble_create(&created, &sim) != CUBBLE_OK) return fail(nullptr);
    int rc = CUBBLE_OK;
    cubble_scalars sc;
    cubble_observables obs;
    if ((rc = cubble_get_scalars(sim, &sc)) != CUBBLE_OK) return fail(sim);
    const double *L = sc.interval;  // refreshed below, after any rescale
    // timeScalingFactor = K / avgRad^2 (src/simulator.cu:557-559)
    const double time_scale = cfg.k_param / (cfg.r_mean * cfg.r_mean);
    // the AB2 history of init (src/simulator.cu:739-787)
    std::printf("Calculating initial velocities for Adams-Bashforth-Moulton\n");
    if (cubble_bootstrap_velocities(sim)) return fail(sim);
    const int64_t ssteps = rs.stabilization_steps;
    if (ssteps > 0 && five_rounds(sim, ssteps, "creation")) return fail(sim);
    if (lattice && (scale_box(sim, cfg) || (ssteps > 0 && five_rounds(sim, ssteps, "scaling"))))
        return fail(sim);
    if (cubble_get_scalars(sim, &sc)) return fail(sim);
    int64_t row_searches0 = sc.searches;
    if (ssteps > 0 &&
        stabilization_phase(sim, ssteps, rs.max_delta_energy, time_scale, row_searches0))
        return fail(sim);
    // integration starts from zero time, wrap counters and errors, with the
    // energy of the stabilised foam (src/simulator.cu:938-955)
    if (cubble_reset_clock(sim, 1) || cubble_reset_print_window(sim) ||
        cubble_observe(sim, &obs, nullptr) || cubble_get_scalars(sim, &sc))
        return fail(sim);
    double e1 = obs.total_energy;
    const int64_t steps0 = sc.steps;  // run.maxSteps counts integration steps
    const double min_interval =
        0.5 * (3 == sc.dim ? std::fmin(L[0], std::fmin(L[1], L[2])) : std::fmin(L[0], L[1]));

    Snapshots snaps(rs.snapshot_name);
    int64_t num_snaps = 0;
    if (snap_freq > 0.0) {
        if (snaps.start(sim, sc) || snaps.write(sim, sc)) return fail(sim);
        ++num_snaps;
    }

    std::printf("\n===========\nIntegration\n===========\n");
    std::printf("%-5s %-8s %-6s %-11s %9s %10s %-6s %-9s %-11s %-11s %-11s \n", "T", "phi", "R",
                "dE", "#b   ", "#pairs", "#steps", "#searches", "min ts", "max ts", "avg ts");

    const cubble_integrate_params ip = {1, cfg.flow, 1};
    cubble_step_result res;
    int64_t printed = 1;
    int64_t row_steps0 = sc.steps;
    // the reference counts the first row's steps from 0 at its first step
    // (the counter is incremented after the print check, src/simulator.cu:1134),
    // so its first row reports one step fewer and averages over that count
    bool first_row = true;
    for (;;) {
        const SplitTime t_print = split((double)printed / time_scale);
        SplitTime t_snap = {~0ull, 0.0}, next = t_print;
        if (snap_freq > 0.0) {
            t_snap = split((double)num_snaps / (snap_freq * time_scale));
            if (earlier(t_snap, next)) next = t_snap;
        }
        const cubble_stop_rules rules = {cfg.num_end, 1, min_interval, next.i, next.f};
        int64_t chunk = kChunk;
        if (cfg.max_steps > 0) chunk = std::min<int64_t>(chunk, cfg.max_steps - (sc.steps - steps0));
        if ((rc = cubble_set_stop_rules(sim, &rules)) ||
            (rc = cubble_run_steps(sim, &ip, chunk, &res)) || (rc = cubble_get_scalars(sim, &sc)))
            break;
        const bool more = sc.count > cfg.num_end && sc.max_r < min_interval &&
                          (cfg.max_steps <= 0 || sc.steps - steps0 < cfg.max_steps);
        const bool print = reached(sc, t_print);
        if (print) {
            if ((rc = cubble_observe(sim, &obs, nullptr))) break;
            const double e2 = obs.total_energy;
            double de = std::abs(e2 - e1);
            if (de > 0.0) de *= 2.0 / (e2 + e1);
            const double rel_rad = obs.mean_r / cfg.r_mean;
            const double *v = obs.mean_v;
            std::ofstream results("results.dat", std::ios_base::app);
            if (results.is_open())
                results << printed << " " << rel_rad << " " << sc.count << " " << obs.mean_path
                        << " " << e2 << " " << de << " " << v[0] << " " << v[1] << " " << v[2]
                        << " " << std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) << " "
                        << v[3] << "\n";
            else
                std::printf("Couldn't open file stream to append results to!\n");
            const int64_t nsteps = sc.steps - row_steps0 - (first_row ? 1 : 0);
            std::printf("%-5d %-#8.6g %-#6.4g %-9.5e %9d %10d %6ld %-9ld %-9.5e %-9.5e %-9.5e \n",
                        (int)printed, obs.total_volume / obs.box_volume, rel_rad, de, sc.count,
                        sc.num_pairs, (long)nsteps, (long)(sc.searches - row_searches0),
                        sc.ts_min, sc.ts_max, sc.ts_sum / (double)nsteps);
            ++printed;
            e1 = e2;
            row_steps0 = sc.steps;
            row_searches0 = sc.searches;
            first_row = false;
        }
        if (snap_freq > 0.0 && reached(sc, t_snap)) {
            if ((rc = snaps.write(sim, sc))) break;
            ++num_snaps;
        }
        if (print && (rc = cubble_reset_print_window(sim))) break;
        if (!more) break;
    }
    if (rc != CUBBLE_OK) return fail(sim);
    if (sc.count <= cfg.num_end)
        std::printf("Stopping simulation, since the number of bubbles left in the simulation (%d) "
                    "is less than or equal to the specified minimum (%d)\n",
                    sc.count, cfg.num_end);
    else if (sc.max_r > min_interval)
        std::printf("Stopping simulation, since the radius of the largest bubble (%g) is greater "
                    "than the simulation box (%g, %g, %g)\n",
                    sc.max_r, L[0], L[1], L[2]);
    else if (cfg.max_steps > 0 && sc.steps - steps0 >= cfg.max_steps)
        std::printf("Stopping simulation after %lld steps (run.maxSteps)\n",
                    (long long)(sc.steps - steps0));
    else
        std::printf("Stopping simulation for an unknown reason...\n");
    if (snap_freq > 0.0 && snaps.write(sim, sc)) return fail(sim);
    std::printf("Cleaning up...\n");
    cubble_destroy(sim);
    std::printf("Done\n");
    return EXIT_SUCCESS;
}
