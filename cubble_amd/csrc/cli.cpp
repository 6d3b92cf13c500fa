// cli.cpp -- drop-in command line entry, cubble_run(argc, argv), and the
// one-line main of cubble-cli.  Mirrors cubble::run (reference
// src/simulator.cu:1161-1186) and the integration loop of simulate()
// (src/simulator.cu:958-1157): argument check, device check, integrate until
// the bubble count reaches numEnd or the largest bubble reaches half the box,
// printing the reference's integration table once per scaled time unit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "cubble_amd.h"

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
    if (cubble_config_from_json(argv[1], &cfg) != CUBBLE_OK) {
        std::fprintf(stderr, "error: %s\n", cubble_last_error(nullptr));
        return EXIT_FAILURE;
    }
    cubble_sim *sim = nullptr;
    if (cubble_create(&cfg, &sim) != CUBBLE_OK) {
        std::fprintf(stderr, "error: %s\n", cubble_last_error(nullptr));
        return EXIT_FAILURE;
    }
    cubble_scalars sc;
    cubble_get_scalars(sim, &sc);
    const double min_interval =
        0.5 * (3 == sc.dim ? std::fmin(sc.interval[0], std::fmin(sc.interval[1], sc.interval[2]))
                           : std::fmin(sc.interval[0], sc.interval[1]));
    const double time_scale = cfg.k_param / (cfg.r_mean * cfg.r_mean);
    std::printf("%-5s %-8s %-6s %9s %10s %-6s %-9s %-11s\n", "T", "phi", "R", "#b   ", "#pairs",
                "#steps", "#searches", "ts");
    cubble_integrate_params ip = {1, 0, 1};
    cubble_step_result res;
    int rc = CUBBLE_OK;
    int64_t printed = 1, last_steps = 0, last_searches = 0;
    const double vbox = sc.interval[0] * sc.interval[1] * (3 == sc.dim ? sc.interval[2] : 1.0);
    for (;;) {
        rc = cubble_run_steps(sim, &ip, 100, &res);
        if (rc != CUBBLE_OK) break;
        const bool more = res.count > cfg.num_end && res.max_radius < min_interval &&
                          (cfg.max_steps <= 0 || res.steps < cfg.max_steps);
        if (res.time * time_scale >= (double)printed || !more) {
            std::printf("%-5lld %-#8.6g %-#6.4g %9d %10d %6lld %-9lld %-9.5e\n", (long long)printed,
                        0.0 * vbox + cfg.phi, 1.0, res.count, res.num_pairs,
                        (long long)(res.steps - last_steps),
                        (long long)(res.searches - last_searches), res.dt);
            ++printed;
            last_steps = res.steps;
            last_searches = res.searches;
        }
        if (!more) break;
    }
    if (rc != CUBBLE_OK) {
        std::fprintf(stderr, "error: %s\n", cubble_last_error(sim));
        cubble_destroy(sim);
        return EXIT_FAILURE;
    }
    if (res.count <= cfg.num_end)
        std::printf("Stopping simulation, since the number of bubbles left in the simulation (%d) "
                    "is less than or equal to the specified minimum (%d)\n",
                    res.count, cfg.num_end);
    else if (res.max_radius >= min_interval)
        std::printf("Stopping simulation, since the radius of the largest bubble (%g) is greater "
                    "than half the simulation box\n",
                    res.max_radius);
    else
        std::printf("Stopping simulation after %lld steps\n", (long long)res.steps);
    cubble_destroy(sim);
    std::printf("Done\n");
    return EXIT_SUCCESS;
}
