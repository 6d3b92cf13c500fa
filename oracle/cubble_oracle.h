/*
 * cubble_oracle.h -- C interface of the CPU FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a single-threaded C++ restatement
 * of the reference simulator's per-timestep path (KJLankinen/cubble,
 * src/kernels.cu and src/simulator.cu).  It exists to check the HIP product
 * library and to time a CPU baseline.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it; the product library never
 * links, calls or falls back to it.
 *
 * Array slots: component index 0 = x, 1 = y, 2 = z, 3 = r for every
 * per-bubble vector quantity (z unused in 2D).
 */
#ifndef CUBBLE_ORACLE_H
#define CUBBLE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_consts {
    int32_t dim;            /* 2 or 3 */
    int32_t wall[3];        /* 1 = wall on that axis, 0 = periodic */
    double lbb[3];          /* lower back bottom corner */
    double tfr[3];          /* top front right corner */
    double interval[3];     /* tfr - lbb */
    double wall_drag;       /* box.wall.drag */
    double skin;            /* 0.3 * radius.mean */
    double min_rad;         /* 0.1 * radius.mean */
    double f0;              /* (float)sigma * radius.mean / mu */
    double k_param;         /* constants.K */
    double kappa;           /* constants.kappa */
    double avg_area_in;     /* mean initial surface area */
    double vol_mul;         /* volume multiplier M */
    double tol;             /* errorTolerance */
    int32_t cell_cap;       /* maximum number of cells in the cell grid */
    double flow_lbb[3];     /* imposed-flow box, box coordinates */
    double flow_tfr[3];
    double flow_vel[3];     /* imposed velocity (already times f0) */
} ora_consts;

typedef struct ora_scalars {
    int32_t count;
    int32_t cell_dim[3];
    double dt;
    uint64_t time_int;
    double time_frac;
    double max_r_cell;      /* max radius used for the cell grid */
    int64_t steps;          /* accepted integrate() calls */
    int64_t attempts;       /* step() calls */
    int64_t searches;
    int64_t removals;
    double bubble_steps;    /* sum of live count over accepted steps */
    /* outputs of the last step() attempt */
    double s_a, s_ar, s_oa, s_oar, s_v;
    double max_err, max_r, max_exp;
    int32_t n_del;
    int32_t accepted;
    int32_t need_search;
    int32_t num_pairs;      /* half-list pair count */
} ora_scalars;

typedef struct ora_state {
    double *x[4];           /* x, y, z, r (r in slot 3) */
    double *v[4];           /* dxdt, dydt, dzdt, drdt */
    double *vo[4];          /* old velocities */
    double *saved[4];       /* positions / radii at last search */
    double *path;
    double *err;
    int32_t *wc[3];         /* wrap counters */
    int32_t *idx;           /* original bubble index */
    int32_t *nbr_off;       /* CSR offsets, count + 1 entries */
    int32_t *nbr;           /* CSR full neighbour lists (both directions) */
    int64_t nbr_cap;        /* capacity of nbr on get, length on set */
    int32_t *deg;           /* list length at the last search (imposed flow);
                               set: NULL = the current list lengths */
} ora_state;

typedef struct ora_attempt_out {
    double *xc[4];          /* corrected positions / radius of the attempt */
    double *vp[4];          /* predicted velocities (drdt final incl. gas) */
    double *ov;             /* per-bubble overlap area sum */
    double *path_new;
    uint8_t *del;           /* 1 = marked for deletion */
} ora_attempt_out;

typedef struct ora_init_params {
    int32_t dim;
    int32_t count;
    int32_t radius_dist;    /* 0 = clipped normal, 1 = log-normal, 2 = |normal| (+ minRad at or below minRad) */
    double r_mean;          /* normal mean, or log-normal median */
    double r_std;           /* normal sd, or log-normal sigma */
    double r_lo, r_hi;      /* clip bounds (normal only) */
    double phi;
    double rel_dim[3];
    int32_t wall[3];
    double wall_drag;
    double avg_rad;         /* radius.mean for skin / minRad / f0 */
    double mu, sigma, kappa, k_param, tol, dt0;
    double cell_cap_factor; /* cell capacity = factor * initial cells */
    uint64_t seed;
} ora_init_params;

void *ora_create(const ora_consts *c, int32_t count);
void ora_destroy(void *f);
void ora_get_consts(void *f, ora_consts *c);
void ora_set_state(void *f, const ora_state *s, const ora_scalars *sc);
int ora_get_state(void *f, ora_state *s, ora_scalars *sc);
void ora_get_attempt(void *f, ora_attempt_out *o);

/* One step() attempt (predict, pair, correct, dt control); no commit. */
int ora_attempt(void *f, int gas, int path, int flow);
/* integrate(): attempts until accepted, commit, removal, search. */
int ora_integrate(void *f, int gas, int path, int flow);
/* integrate() repeated n times; returns seconds of wall time. */
double ora_run(void *f, int gas, int path, int flow, int64_t n);
void ora_search(void *f);
void ora_remove(void *f);
/* O(N^2) pair set (i < j) from the current state, ascending. */
int64_t ora_brute_pairs(void *f, int32_t *out, int64_t cap);
/* Synthetic initial foam (positions, radii, constants) plus first search. */
void *ora_init_synthetic(const ora_init_params *p);

/* Observables of the integration loop's printed rows and snapshots:
 * out[0..7] = total energy, total volume, then the sums of r, path, dxdt,
 * dydt, dzdt, drdt; energy (count doubles or NULL) = each bubble's energy. */
void ora_observe(void *f, double out[8], double *energy);

/* The AB2 velocity bootstrap of init: old velocities = pair velocities
 * after one Euler step of the pair velocities (no gas, flow or walls). */
void ora_bootstrap(void *f);

/* The box rescale of init to gas fraction phi (box t * rel), given the
 * bubbles' total volume; then a search. */
void ora_rescale(void *f, double phi, const double rel[3], double vol);
/* The wall term of the corrector (src/kernels.cu:546-593) on its own:
 * v (x, y, z) is updated in place for a bubble at pos with radius rad. */
void ora_wall_velocity(const ora_consts *c, const double pos[3], double rad, double v[3]);
/* Deterministic cube root used by the volume fix. */
double ora_cbrt(double a);
/* Raw outputs of the project RNG, for known-answer tests: SplitMix64 from
 * `seed`, and xoshiro256** from an explicit state. */
void ora_splitmix64(uint64_t seed, int32_t n, uint64_t *out);
void ora_xoshiro256ss(const uint64_t state[4], int32_t n, uint64_t *out);
void ora_set_threads(int n);
int ora_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
