/*
 * cubble_amd.h -- C ABI of the MI355X-native cubble per-timestep path.
 *
 * The reference (KJLankinen/cubble) exports exactly one linkable symbol,
 * `int32_t cubble::run(int32_t argc, char **argv)` (src/simulator.cu:1160-1187),
 * and keeps its step interface -- step(), integrate(), searchNeighbors(),
 * removeBubbles() -- in an anonymous namespace (src/simulator.cu:36-337).
 * This header exposes that same interface as plain C: an opaque simulation
 * handle owning device memory and a HIP stream, plain pointers and sizes,
 * integer status codes, no exceptions and no torch types.
 *
 * Threading: one host thread per handle at a time.  Separate handles may be
 * driven concurrently from separate threads; each owns its own stream.
 *
 * Array slot convention used by every state view: component 0 = x, 1 = y,
 * 2 = z, 3 = r.  Per-bubble arrays are `count` long; z arrays are ignored
 * (and may be NULL) in 2D.  Any NULL pointer in a view is skipped.
 */
#ifndef CUBBLE_AMD_H
#define CUBBLE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CUBBLE_OK 0
#define CUBBLE_E_INVALID (-1)     /* bad argument or configuration            */
#define CUBBLE_E_HIP (-2)         /* HIP runtime error                         */
#define CUBBLE_E_NOMEM (-3)       /* device memory exhausted                   */
#define CUBBLE_E_IO (-4)          /* input file missing or malformed           */
#define CUBBLE_E_UNSUPPORTED (-5) /* feature outside this build (imposed flow) */
#define CUBBLE_E_CAPACITY (-6)    /* a capacity limit was exceeded             */
#define CUBBLE_E_STATE (-7)       /* call not valid in the handle's state      */

/* ---- precision modes ----------------------------------------------------- */
#define CUBBLE_PREC_FP64 0  /* everything in FP64                               */
#define CUBBLE_PREC_MIXED 1 /* pair forces and gas fluxes in FP32 on relative
                               coordinates (32-bit fixed-point positions);
                               velocities, saved positions, per-bubble error
                               and overlap stored in FP32; positions, radii,
                               the corrector, error estimate and volume fix
                               in FP64                                           */

typedef struct cubble_sim cubble_sim;

/* Mirrors the reference's IntegrationParams flags
 * (src/data_definitions.h:227-236).  use_flow adds the imposed flow of the
 * handle's configuration (flow.*; addFlowVelocity, src/kernels.cu:497-544). */
typedef struct cubble_integrate_params {
    int32_t use_gas_exchange;
    int32_t use_flow;
    int32_t increment_path;
} cubble_integrate_params;

/* Mirrors the outputs the reference's step() leaves in IntegrationParams and
 * HostData (src/simulator.cu:219-246, :303-321). */
typedef struct cubble_step_result {
    int32_t error_too_large;    /* last attempt rejected (ip.errorTooLarge)  */
    int32_t num_to_be_deleted;  /* bubbles marked r <= minRad in last attempt */
    int32_t count;              /* live bubbles                               */
    int32_t num_pairs;          /* half-list pairs (i < j)                    */
    int32_t need_search;        /* last accepted attempt triggered a search   */
    int32_t halted;             /* device loop stopped (limit or capacity)    */
    double max_error;
    double max_radius;
    double max_expansion;
    double dt;                  /* time step for the next attempt             */
    double time;                /* timeInteger + timeFraction                 */
    int64_t steps;              /* accepted integrate() calls                 */
    int64_t attempts;           /* step() calls                               */
    int64_t searches;
    int64_t removals;
    double bubble_steps;        /* sum of live count over accepted steps      */
} cubble_step_result;

/* Programmatic configuration of a synthetic initial foam.  Field meanings
 * follow input_parameters.json (reference input_parameters.json:1-130) plus
 * the initial-distribution keys this build adds. */
typedef struct cubble_config {
    int32_t dim;              /* box.dimensionality: 2 or 3                  */
    int32_t num_start;        /* bubbles.numStart                            */
    int32_t num_end;          /* bubbles.numEnd (CLI stop rule)              */
    int32_t radius_dist;      /* 0 = normal clipped to [r_lo, r_hi],
                                 1 = log-normal (median r_mean, sigma r_std),
                                 2 = the reference's rule: |N(r_mean, r_std)|,
                                 plus minRad (0.1 r_mean) when at or below
                                 minRad (src/kernels.cu:1051-1053); the JSON
                                 reader's default when the file has no
                                 "init" object                              */
    double r_mean;            /* bubbles.radius.mean                         */
    double r_std;             /* bubbles.radius.std                          */
    double r_lo, r_hi;        /* clip bounds for the normal distribution     */
    double rel_dim[3];        /* box.relativeDimensions                      */
    int32_t wall[3];          /* box.wall.{x,y,z}                            */
    double wall_drag;         /* box.wall.drag                               */
    double phi, mu, sigma, kappa, k_param; /* constants.*.value              */
    double error_tolerance;   /* errorTolerance.value                        */
    double dt0;               /* initial time step (reference: 1e-4)         */
    uint64_t seed;            /* rngSeed.value                               */
    int32_t precision;        /* CUBBLE_PREC_*                               */
    int32_t nbr_cap;          /* initial neighbour capacity per bubble (0 = default) */
    double cell_cap_factor;   /* cell capacity / initial cells (0 = default 2) */
    int64_t max_steps;        /* CLI stop rule: accepted steps (0 = no limit) */
    int32_t flow;             /* flow.impose: the CLI integrates with use_flow */
    double flow_vel[3];       /* flow.velocity, in units of f0 = sigma R / mu  */
    double flow_lbb[3];       /* flow.lbb, relative to the box (0..1)          */
    double flow_tfr[3];       /* flow.tfr, relative to the box (0..1)          */
    int32_t placement;        /* 0 = uniform positions in the box; 1 = the
                                 reference's lattice (spacing 2 R, a random
                                 offset of up to R per bubble, then scaled to
                                 the box; src/simulator.cu:576-606,
                                 src/kernels.cu:1026-1064); 2 = that lattice
                                 in its own box (edge 2 R per site), for
                                 cubble_rescale_box later.  With 1 or 2,
                                 num_start must be the lattice count (the JSON
                                 reader rounds numStart up to it, as the
                                 reference)                                    */
} cubble_config;

/* State view: caller-owned host buffers, filled by cubble_get_state or read
 * by cubble_set_state.  Neighbour lists are CSR, full (both directions),
 * ascending per bubble; the reference's half list is the subset j > i. */
typedef struct cubble_state_view {
    double *x[4];             /* positions + radius                          */
    double *v[4];             /* dxdt .. drdt                                */
    double *vo[4];            /* dxdto .. drdto                              */
    double *saved[4];         /* savedX .. savedR                            */
    double *path;
    double *error;
    int32_t *wrap[3];         /* wrapCountX .. Z                             */
    int32_t *index;           /* original bubble index                       */
    int32_t *nbr_off;         /* count + 1 offsets                           */
    int32_t *nbr;             /* neighbour list                              */
    int64_t nbr_cap;          /* length of nbr (get: capacity, set: length)  */
    int32_t *degree;          /* list length at the last search, which the
                                 imposed flow divides by (the reference's
                                 numNeighbors: a removal moves it with the
                                 bubble but does not reduce it); set: NULL =
                                 the current list lengths                      */
} cubble_state_view;

/* Outputs of the last step() attempt (the reference's predicted buffers). */
typedef struct cubble_attempt_view {
    double *xc[4];            /* corrected positions / radius                */
    double *vp[4];            /* predicted velocities (drdt incl. gas)       */
    double *overlap;          /* per-bubble overlap area sum                 */
    double *path_new;
    uint8_t *deleted;         /* 1 = marked for deletion                     */
} cubble_attempt_view;

/* Scalars of a handle (constants and dynamic state). */
typedef struct cubble_scalars {
    int32_t dim;
    int32_t count;
    int32_t capacity;
    int32_t nbr_cap;
    int32_t cell_cap;
    int32_t cell_dim[3];
    int32_t wall[3];
    int32_t precision;
    double lbb[3], tfr[3], interval[3];
    double wall_drag, skin, min_rad, f0, k_param, kappa, avg_area_in, vol_mul, tol;
    double dt;
    uint64_t time_int;
    double time_frac;
    double max_r_cell;
    int64_t steps, attempts, searches, removals;
    double bubble_steps;
    double s_a, s_ar, s_oa, s_oar, s_v, max_err, max_r, max_exp;
    int32_t n_del, accepted, need_search, num_pairs;
    double ts_min, ts_max;    /* extremes of the accepted steps' time step
                                 since the last cubble_reset_print_window
                                 (empty window: 9999999.9 and -1.0, the
                                 reference's, src/simulator.cu:980-982)       */
    double ts_sum;            /* their sum (src/simulator.cu:1021-1023)       */
    double flow_lbb[3], flow_tfr[3]; /* flow box, box coordinates
                                        (src/simulator.cu:812-817)             */
    double flow_vel[3];       /* imposed velocity times f0 (:561-565)          */
} cubble_scalars;

/* ---- lifetime ------------------------------------------------------------ */

/* Replaces init() (src/simulator.cu:527-956) for a synthetic foam: draws the
 * radii and positions on the host with the project RNG, uploads them, and
 * runs the first neighbour search (src/simulator.cu:725).  Uses device 0 of
 * the current HIP context. */
int cubble_create(const cubble_config *cfg, cubble_sim **out);

/* Same, reading the reference's input_parameters.json schema
 * (src/simulator.cu:538-575; keys listed in input_parameters.json). */
int cubble_create_from_json(const char *json_path, cubble_sim **out);

/* Parses input_parameters.json into a config without touching the GPU. */
int cubble_config_from_json(const char *json_path, cubble_config *out);

/* Fills `cfg` with the defaults of input_parameters.json. */
void cubble_config_default(cubble_config *cfg);

/* Frees device memory and the stream (replaces end(), src/simulator.cu:514-525). */
void cubble_destroy(cubble_sim *sim);

/* Message of the last failing call on `sim` (or, when sim is NULL, of the
 * last failing create or call without a handle).  Valid until the next call
 * on the same handle. */
const char *cubble_last_error(const cubble_sim *sim);

/* ---- the step interface --------------------------------------------------- */

/* One attempt == step() (src/simulator.cu:192-249): predictor, pairwise
 * velocity and gas flux, corrector, dt control.  Does not commit. */
int cubble_step(cubble_sim *sim, const cubble_integrate_params *ip, cubble_step_result *res);

/* integrate() (src/simulator.cu:251-337): attempts until accepted, commit,
 * removal of dissolved bubbles, neighbour search when the skin is half used. */
int cubble_integrate(cubble_sim *sim, const cubble_integrate_params *ip, cubble_step_result *res);

/* integrate() repeated `num_steps` times, driven entirely on the device (no
 * host synchronisation per step).  Replaces the main loop's integrate calls
 * (src/simulator.cu:1008-1011).  Returns when the steps are done.
 * num_steps = 0 only prepares the device loop for these parameters (captures
 * its graph), so that a following run starts at once. */
int cubble_run_steps(cubble_sim *sim, const cubble_integrate_params *ip, int64_t num_steps,
                     cubble_step_result *res);

/* searchNeighbors() (src/simulator.cu:60-178). */
int cubble_search_neighbors(cubble_sim *sim);

/* removeBubbles() (src/simulator.cu:180-190) using the deletion marks of the
 * last attempt. */
int cubble_remove_bubbles(cubble_sim *sim);

/* ---- observables, stop rules and the print window ------------------------ */

/* What the reference's integration loop reports at a printed row and in a
 * snapshot (src/simulator.cu:1036-1082, :387-512): totalEnergy
 * (src/simulator.cu:39-58, kernel potentialEnergy src/kernels.cu:941-956),
 * totalVolume (src/simulator.cu:362-379), boxVolume (:381-385) and the means
 * of getAvg (:1036-1050), reduced on the device. */
typedef struct cubble_observables {
    double total_energy;      /* sum over bubbles i of the sum over overlapping
                                 neighbours j of (r_i + r_j - |d_ij|)^2      */
    double total_volume;      /* sum of pi r^2 (2-D) or 4/3 pi r^3 (3-D)      */
    double box_volume;
    double mean_r;
    double mean_path;
    double mean_v[4];         /* dxdt, dydt, dzdt, drdt                       */
} cubble_observables;

/* energy (count doubles, or NULL) receives each bubble's energy in the order
 * of cubble_get_state: the snapshot's energy column (src/simulator.cu:388-392). */
int cubble_observe(cubble_sim *sim, cubble_observables *out, double *energy);

/* Stop rules the device loop checks after every accepted step, so that
 * cubble_run_steps returns at the step where the reference's loop acts:
 * its stop condition (src/simulator.cu:1016-1018) and its print and snapshot
 * schedule (:1025-1034, :1110-1122).  A run that meets a rule returns early,
 * with fewer steps done; a rule already met when a run starts ends it after
 * its first step.  Rules stay set until replaced (all zero = none). */
typedef struct cubble_stop_rules {
    int32_t min_count;        /* halt once count <= min_count after a removal
                                 (0 = off)                                    */
    int32_t use_time;         /* 1: halt once time >= time_int + time_frac,
                                 compared as the reference does (integer part
                                 first, then the fraction)                    */
    double max_radius;        /* halt once a step's largest radius >= this
                                 (0 = off)                                    */
    uint64_t time_int;
    double time_frac;
} cubble_stop_rules;
int cubble_set_stop_rules(cubble_sim *sim, const cubble_stop_rules *rules);

/* What the reference resets after each printed row (src/simulator.cu:1099-1106,
 * :1127-1131): the per-bubble error array, and the window of time steps that
 * cubble_scalars.ts_min / ts_max / ts_sum report. */
int cubble_reset_print_window(cubble_sim *sim);

/* The box rescale of the reference's init (src/simulator.cu:793-846): the
 * box becomes t * rel_dim (rel_dim[2] ignored in 2-D) with t chosen so the
 * bubbles' volume fills the fraction phi; every coordinate keeps its relative
 * place, the flow box likewise, the volume multiplier follows the new box,
 * and the neighbour lists are searched again.  For foams created with
 * placement 2 (the lattice in its own box).  CUBBLE_E_INVALID for phi <= 0,
 * CUBBLE_E_STATE for a foam without bubble volume. */
int cubble_rescale_box(cubble_sim *sim, double phi, const double rel_dim[3]);

/* The AB2 velocity bootstrap of the reference's init (src/simulator.cu:
 * 739-787): the pair velocities (no gas exchange, flow or walls) at the
 * positions moved by one Euler step of dt along the pair velocities become
 * the old velocities; the current ones stay zero.  Only before the first
 * attempt (CUBBLE_E_STATE otherwise); without it the history starts at zero. */
int cubble_bootstrap_velocities(cubble_sim *sim);

/* What the reference's init resets before each stabilisation round and
 * before the integration loop (src/simulator.cu:866-867, :938-955): the
 * clock to zero and, with clear_wraps, the wrap counters, so that unwrapped
 * positions and snapshot displacements start from the stabilised foam. */
int cubble_reset_clock(cubble_sim *sim, int32_t clear_wraps);

/* ---- state access (tests, checkpoints) ----------------------------------- */
int cubble_get_scalars(cubble_sim *sim, cubble_scalars *out);
int cubble_get_state(cubble_sim *sim, cubble_state_view *view);
int cubble_get_attempt(cubble_sim *sim, cubble_attempt_view *view);
/* Replaces the whole state (count, arrays, scalars dt/time/max_r_cell/counters). */
int cubble_set_state(cubble_sim *sim, const cubble_state_view *view, const cubble_scalars *sc);

/* ---- launch shapes ------------------------------------------------------- */

/* Launch-shape parameters picked by the autotuner (0 = built-in default).
 * No counterpart in the reference (its launches are fixed at 5120 x 384,
 * src/data_definitions.h:28-29). */
typedef struct cubble_tuning {
    int32_t bubbles_per_thread;  /* bubble-space kernels (0 = from the foam size)   */
    int32_t search_tile;         /* bubbles per neighbour-search workgroup
                                    (0 = 256, >= 32)                        */
    int32_t graph_steps;         /* attempt groups per graph replay (0 = 16)        */
} cubble_tuning;

/* cubble_create with explicit launch shapes. */
int cubble_create_tuned(const cubble_config *cfg, const cubble_tuning *tuning, cubble_sim **out);

/* Autotuner options and report.  The tuner times candidate launch shapes on
 * the handle's own foams, each from the same copy of the device state, keeps
 * the fastest, and restores the state (tuning does not advance the
 * simulation).  With cache_path, shapes are looked up in / stored to a JSON
 * file keyed by device and batch make-up.  No counterpart in the reference. */
typedef struct cubble_autotune_opts {
    int32_t trial_steps;    /* accepted steps timed per candidate (0 = 64)   */
    int32_t repeats;        /* timings per candidate, fastest kept (0 = 2)   */
    int32_t ignore_cache;   /* time the candidates even on a cache hit       */
    const char *cache_path; /* JSON cache file or NULL                       */
} cubble_autotune_opts;

typedef struct cubble_autotune_report {
    cubble_tuning chosen;
    int32_t from_cache;
    int32_t candidates_timed;
    double start_rate;      /* bubble-steps/s with the shapes in use before */
    double best_rate;       /* with the chosen shapes (0 on a cache hit)    */
} cubble_autotune_report;

int cubble_autotune(cubble_sim *sim, const cubble_integrate_params *ip,
                    const cubble_autotune_opts *opts, cubble_autotune_report *report);
/* Apply launch shapes to a live simulation (state untouched). */
int cubble_retune(cubble_sim *sim, const cubble_tuning *tuning);

/* ---- batched foams --------------------------------------------------------- */

/* Many independent foams (2D and 3D mixed, uneven sizes, each with its own
 * constants) advanced together: every kernel of a step is launched once for
 * the whole batch over a block table, the accept / reject, dt, removal and
 * search decisions are taken per foam on the device, and a run of steps is
 * a captured HIP graph replayed without host synchronisation per step.
 * No counterpart in the reference (one foam per process).  Each foam's
 * results are bit-identical to the same foam run alone. */
typedef struct cubble_batch cubble_batch;

/* All configurations must share `precision`. */
int cubble_batch_create(const cubble_config *cfgs, int32_t n, const cubble_tuning *tuning,
                        cubble_batch **out);
int cubble_batch_size(const cubble_batch *b);
/* One accepted integrate() step for every foam. per_foam: n results or NULL. */
int cubble_batch_integrate(cubble_batch *b, const cubble_integrate_params *ip,
                           cubble_step_result *per_foam);
/* `num_steps` accepted steps for every foam, device-driven (0: prepare only,
 * as cubble_run_steps). */
int cubble_batch_run_steps(cubble_batch *b, const cubble_integrate_params *ip, int64_t num_steps,
                           cubble_step_result *per_foam);
int cubble_batch_get_scalars(cubble_batch *b, int32_t k, cubble_scalars *out);
int cubble_batch_get_state(cubble_batch *b, int32_t k, cubble_state_view *view);
const char *cubble_batch_last_error(const cubble_batch *b);
int cubble_batch_autotune(cubble_batch *b, const cubble_integrate_params *ip,
                          const cubble_autotune_opts *opts, cubble_autotune_report *report);
int cubble_batch_retune(cubble_batch *b, const cubble_tuning *tuning);
void cubble_batch_destroy(cubble_batch *b);

/* ---- serving ----------------------------------------------------------------- */

/* A request loop for foam simulations with continuous batching: requests are
 * admitted in arrival order when their device-memory footprint fits the
 * budget (and a resident slot is free), their state stays resident on the
 * GPU, and all resident foams advance together in rounds of `round_steps`
 * accepted steps; requests that arrive while others run are admitted at the
 * next round boundary, and finished requests free their memory.  A request
 * finishes after `max_steps` accepted steps or once its bubble count has
 * fallen to (1 - dissolve_fraction) of the start.  No counterpart in the
 * reference.  Each request's result equals the same foam run alone. */
typedef struct cubble_server cubble_server;

typedef struct cubble_server_opts {
    uint64_t hbm_budget_bytes; /* device memory for resident foams             */
    int32_t max_active;        /* resident foams at once (0 = no limit)         */
    int32_t round_steps;       /* accepted steps per round (0 = 50)             */
    int32_t keep_states;       /* keep final states for cubble_server_get_state */
    int32_t precision;         /* CUBBLE_PREC_* of every request                */
} cubble_server_opts;

typedef struct cubble_request_result {
    int64_t id;
    int32_t status;            /* 0 queued, 1 resident, 2 done, < 0 error code  */
    int32_t start_count;
    int32_t final_count;
    int32_t pad;
    int64_t steps;
    double bubble_steps;
    double arrival_s;          /* seconds from the start of cubble_server_run   */
    double admit_s;
    double finish_s;
} cubble_request_result;

typedef struct cubble_server_stats {
    int64_t requests;
    int64_t rounds;
    int32_t max_resident;
    int32_t pad;
    double wall_s;
    double bubble_steps;
    double throughput;         /* bubble-steps/s over the run                   */
    double p50_latency_s;      /* completion - arrival                          */
    double p99_latency_s;
    double peak_bytes;         /* device memory held by resident foams          */
} cubble_server_stats;

int cubble_server_create(const cubble_server_opts *opts, cubble_server **out);
/* Queue a request arriving `arrival_s` seconds after the run starts.  A
 * request with imposed flow (cfg->flow) is refused with
 * CUBBLE_E_UNSUPPORTED: the resident foams share one set of launch flags. */
int cubble_server_submit(cubble_server *s, const cubble_config *cfg, int64_t max_steps,
                         double dissolve_fraction, double arrival_s, int64_t *id);
/* Serve every queued request (arrivals in real time); returns when all are done. */
int cubble_server_run(cubble_server *s, cubble_server_stats *stats);
int cubble_server_result(cubble_server *s, int64_t id, cubble_request_result *out);
int cubble_server_get_state(cubble_server *s, int64_t id, cubble_state_view *view,
                            cubble_scalars *sc);
const char *cubble_server_last_error(const cubble_server *s);
void cubble_server_destroy(cubble_server *s);

/* Device-memory footprint of one foam of this configuration (admission size). */
int64_t cubble_foam_footprint(const cubble_config *cfg);

/* ---- measurement ---------------------------------------------------------- */

/* Average device time of each kernel of one attempt from the current state,
 * measured with HIP events on the handle's stream around `reps` back-to-back
 * launches of each (predictor; pair kernel; sums + corrector + controller,
 * not committing).  No counterpart in the reference (it has NVTX ranges
 * only, src/simulator.cu:40-363); used by the benchmark's roofline report. */
typedef struct cubble_kernel_times {
    double predict_ms;
    double pair_ms;
    double correct_ms;
    int32_t reps;
} cubble_kernel_times;
int cubble_time_kernels(cubble_sim *sim, const cubble_integrate_params *ip, int32_t reps,
                        cubble_kernel_times *out);
int cubble_batch_time_kernels(cubble_batch *b, const cubble_integrate_params *ip, int32_t reps,
                              cubble_kernel_times *out);

/* The search phase (searchNeighbors, src/simulator.cu:60-178), timed from a
 * copy of the device state that is put back afterwards, each figure the
 * average of `reps` back-to-back launches between one pair of HIP events:
 * a whole search of every foam; the list build alone (lists of the
 * reference's neighborSearch / sortPairs, src/kernels.cu:617-719); the list
 * kernel alone; the reorder (reorganizeByIndex, :721-770); and one attempt
 * group's removal and search kernels when no foam needs them (they return
 * at once).  No counterpart in the reference (it has NVTX ranges only). */
typedef struct cubble_search_times {
    double search_ms;
    double lists_ms;
    double list_kernel_ms;
    double reorder_ms;
    double idle_ms;
    double bubbles;           /* bubbles searched (every foam)                */
    double pairs;             /* half-list pairs the search finds              */
    int32_t reps;
} cubble_search_times;
int cubble_time_search(cubble_sim *sim, int32_t reps, cubble_search_times *out);
int cubble_batch_time_search(cubble_batch *b, int32_t reps, cubble_search_times *out);

/* ---- the pair search on caller-owned device arrays ---------------------- */

/* A box: dimensionality, walls (1 = wall, 0 = periodic) and corners. */
typedef struct cubble_box {
    int32_t dim;
    int32_t wall[3];
    double lbb[3], tfr[3];
} cubble_box;

/* The pair output of the reference's searchNeighbors (src/simulator.cu:
 * 60-178; neighborSearch / comparePair and sortPairs, src/kernels.cu:617-711)
 * for bubbles the caller holds on the device: x, y, z (ignored and may be
 * NULL in 2-D) and r, `count` FP64 values each, positions within one box
 * length of the box and radii positive.  Writes every pair i < j with
 * |wrap(x_i - x_j)|^2 < (r_i + r_j + skin)^2 -- the reference's Pairs {i, j},
 * grouped by first index, each group ascending in j -- to the device arrays
 * pair_i / pair_j, and the number of pairs to the device int64 *num_pairs.
 * With more pairs than pair_cap, nothing is written beyond pair_cap: the
 * groups that fit entirely are complete, the group at the boundary holds
 * some of its pairs, and *num_pairs tells how much room a second call
 * needs (the reference writes past its capacity, src/kernels.cu:630-632).  Enqueued on `stream` (a hipStream_t; NULL = the null stream);
 * the host does not wait.  The scratch (about 40 bytes per bubble) is taken
 * with hipMallocAsync and returned with hipFreeAsync on that stream.
 * CUBBLE_E_INVALID for bad arguments; the message is cubble_last_error(NULL). */
int cubble_find_pairs(const double *x, const double *y, const double *z, const double *r,
                      int32_t count, const cubble_box *box, double skin, int32_t *pair_i,
                      int32_t *pair_j, int64_t pair_cap, int64_t *num_pairs, void *stream);

/* cubble_find_pairs without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops that must not call hipMalloc).  The
 * scratch is a fixed device array of CUBBLE_FIND_PAIRS_NOALLOC_SCRATCH_MIB
 * MiB that every device the library runs on holds until the process ends,
 * enough for CUBBLE_FIND_PAIRS_NOALLOC_MAX bubbles; a larger count returns
 * CUBBLE_E_CAPACITY before any work.  All calls on one device share that
 * device's array: calls that overlap on one device (other streams, other
 * host threads) race and must be serialised by the caller.  Same arguments,
 * results and errors as cubble_find_pairs otherwise. */
#define CUBBLE_FIND_PAIRS_NOALLOC_SCRATCH_MIB 32
#define CUBBLE_FIND_PAIRS_NOALLOC_MAX 800000
int cubble_find_pairs_noalloc(const double *x, const double *y, const double *z, const double *r,
                              int32_t count, const cubble_box *box, double skin, int32_t *pair_i,
                              int32_t *pair_j, int64_t pair_cap, int64_t *num_pairs,
                              void *stream);

/* ---- host-only helpers (no GPU needed) ----------------------------------- */

/* Synthetic initial foam on the host: the exact arrays cubble_create uploads
 * (before the first search).  x4 holds 4 * count doubles (x, y, z, r rows);
 * out_sc receives the derived constants (box, skin, minRad, f0, mean area,
 * volume multiplier, cell capacity). */
int cubble_generate_foam(const cubble_config *cfg, double *x4, cubble_scalars *out_sc);

/* Drop-in command line entry == cubble::run (src/simulator.cu:1161-1186). */
int cubble_run(int argc, char **argv);

/* Version string of the library. */
const char *cubble_version(void);

#ifdef __cplusplus
}
#endif
#endif
