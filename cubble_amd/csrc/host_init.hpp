// host_init.hpp -- host-side configuration, JSON schema reader and the
// synthetic initial foam (project RNG).  No GPU calls.
#pragma once

#include <string>
#include <vector>

#include "cubble_amd.h"
#include "foam.hpp"

namespace cubble_amd {

struct InitOut {
    Consts c;
    double max_r;
    int32_t cell_cap;
};

// Initial cell grid capacity rule shared with the search kernels.
void host_cell_dims(const Consts &c, double max_r, long long cap, int cd[3]);

// Radii and positions of a synthetic foam; x4 receives 4 rows of `count`
// doubles (x, y, z, r).
int generate_foam(const cubble_config &cfg, std::vector<double> &x4, InitOut &out,
                  std::string &err);

// Bubbles per axis of the reference's initial lattice for `num_start`
// bubbles (src/simulator.cu:576-600, in its single-precision arithmetic); the
// lattice count is their product.
void lattice_dims(const cubble_config &cfg, int32_t bpd[3]);

// input_parameters.json (reference schema) -> config.
int parse_config_json(const char *path, cubble_config &cfg, std::string &err);

// Settings of the command line's run that cubble_config does not hold:
// snapShot.{frequency, filename} and stabilization.{steps, maxDeltaEnergy}
// of input_parameters.json (reference src/simulator.cu:568-569, :544,
// :875).  Absent keys: no snapshots, no stabilisation phase.
struct RunSettings {
    double snapshot_frequency = 0.0;
    std::string snapshot_name = "snapshot";
    int64_t stabilization_steps = 0;
    double max_delta_energy = 0.0;
};
int parse_run_json(const char *path, RunSettings &rs, std::string &err);

void config_default(cubble_config &cfg);

// Launch shapes chosen by the autotuner, kept in a JSON file keyed by the
// device and the batch's make-up: {key: {bubbles_per_thread, search_tile,
// graph_steps, rate}}.  Lookup returns false when the file or key is absent.
bool tuning_cache_lookup(const char *path, const std::string &key, int shape[3]);
int tuning_cache_store(const char *path, const std::string &key, const int shape[3], double rate,
                       std::string &err);

}  // namespace cubble_amd
