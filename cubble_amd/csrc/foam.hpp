// foam.hpp -- device-resident foam header shared by the HIP kernels and the
// host runtime.
//
// One `Foam` record per simulated foam lives in device memory.  It holds the
// constants of the foam, the *role* pointers that say which physical buffer
// currently plays x, x_predicted, dxdt, dxdt_old ... (the reference swaps
// these on the host, src/simulator.cu:96-147 and :258-301; here the device
// control code swaps them so no host round trip is needed), the scratch
// arrays of the step, and the dynamic scalars (dt, time, count, flags).
//
// Every kernel is launched over a block table: entry b = (foam, local block),
// so the same kernels drive one foam or a batch of foams of uneven size.
#pragma once

#include <stdint.h>

#include "cubble_amd.h"  // CUBBLE_PREC_* (Foam::precision)

namespace cubble_amd {

constexpr int kBlock = 256;         // threads per block, bubble- and cell-space
constexpr int kMaxBlocks = 2048;    // bubble-space blocks per foam (partials)
constexpr int kNumXcd = 8;          // MI355X: 8 XCDs, workgroups round-robin
constexpr int kPartD = 8;           // doubles per block partial record
constexpr int kPartI = 4;           // ints per block partial record

// reference src/util.cuh:63
constexpr double kPi = 3.1415926535897932384626433832795028841971693993;

struct Consts {
    int32_t dim;
    int32_t wall[3];
    double lbb[3], tfr[3], interval[3];
    double wall_drag, skin, min_rad, f0, k_param, kappa, avg_area_in, vol_mul, tol;
    // imposed flow: the box [flow_lbb, flow_tfr] and the velocity (f0 units)
    double flow_lbb[3], flow_tfr[3], flow_vel[3];
};

// Role pointers: which buffer currently plays each part of the state.
// Velocity buffers hold FP64 in FP64 mode and FP32 in mixed precision, where
// every velocity is an FP32 result of the pair kernel (plus the corrector's
// wall and gas terms, rounded once); kernels read them through field<V>().
struct Roles {
    double *x[4];        // current x, y, z, r
    double *xp[4];       // predicted, then corrected (attempt output)
    void *v[4];          // current velocities
    void *vp[4];         // predicted velocities (attempt output)
    void *vo[4];         // old velocities
    void *vs[4];         // spare (search reorder target)
    void *saved[4];      // positions at the last search (FP32 in mixed precision)
    double *path, *path_new;
    void *err, *err_alt; // per-bubble error (FP32 in mixed precision)
    int32_t *wc[3], *wc_alt[3];
    int32_t *idx, *idx_alt;
};

// One bubble's position and radius as a 32-byte record: a neighbour gather
// touches one cache line instead of four.
struct alignas(32) D4 {
    double x, y, z, r;
};

// Mixed precision only: one bubble's predicted position in 32-bit fixed
// point over the box and its radius in FP32, 16 bytes, so a neighbour gather
// is one 16-byte load.  Periodic axes map the box [lb, lb + L) onto the whole
// 32-bit circle, so the wrapping difference of two codes is the minimum-image
// difference; wall axes map [lb, lb + L] onto [0, 2^30].
struct alignas(16) Q4 {
    int32_t x, y, z;
    float r;
};

struct Scratch {
    D4 *p4;              // predicted x, y, z, r of the current attempt (FP64 mode; else null)
    Q4 *q4;              // the same in fixed point + FP32 (mixed precision; else null)
    D4 *x4;              // current x, y, z, r in cell order, written by the search
    void *ov;            // per-bubble overlap area sum (FP32 in mixed precision)
    uint8_t *del;        // deletion mark of the last attempt
    uint8_t *dirty;      // list needs filtering after a removal (all zero otherwise)
    int32_t *nn;         // neighbour count
    int32_t *deg;        // neighbour count at the last search, moved with the
                         // bubble by a removal but not reduced (imposed flow)
    int32_t *nbr;        // neighbour lists, column-major [nbr_cap][cap]
    int32_t *cell_of;    // cell of each bubble (old order)
    int32_t *cell_count;
    int32_t *cell_start; // num_cells + 1
    int32_t *cell_cursor;
    int32_t *members;    // unsorted cell segments
    double *part;        // bubble-space block partials [nblk][kPartD]
    int32_t *part_i;     // bubble-space block partials [nblk][kPartI]
    int32_t *cpart;      // cell-space block partials (cell-count sums)
    int32_t *cpart_i;    // neighbour-search tile partials [sblk][kPartI]
    int32_t *del_list;   // sorted deleted indices
    int32_t *plan_src;   // source (tail survivor) of hole k
    int32_t *moved_to;   // new slot of tail survivor (old index - new count)
};

struct Dyn {
    int32_t count;
    int32_t cell_dim[3];
    int32_t num_cells;
    int32_t accepted;
    int32_t need_search;
    int32_t need_remove;
    int32_t halt;        // device loop stop (limit reached or capacity error)
    int32_t overflow;    // neighbour list capacity exceeded
    int32_t n_del;
    int32_t n_holes;
    int32_t num_pairs;
    int32_t max_nn;
    int32_t count_stop;  // halt once count <= count_stop after a removal (0 = off)
    int32_t failed;      // the host could not give the foam what it needs (list
                         // growth out of memory): halted for good
    double dt;
    uint64_t time_int;
    double time_frac;
    double max_r_cell;
    int64_t steps, attempts, searches, removals;
    int64_t step_limit;  // halt once steps reaches this (0 = none)
    double bubble_steps;
    double s_a, s_ar, s_oa, s_oar, s_v, max_err, max_r, max_exp;
    double vol_fix;      // multiplier of the pending volume fix
    double search_max_r; // largest radius at the running search (k_s_cells), 0 otherwise
    // stop rules checked after every accepted step (0 = off)
    double r_stop;       // halt once the step's largest radius >= r_stop
    int32_t time_stop;   // halt once time_int + time_frac >= stop_int + stop_frac
    uint64_t stop_int;
    double stop_frac;
    // extremes and sum of the accepted steps' dt since the last window reset
    double ts_min, ts_max, ts_sum;
};

// Empty time-step window (reference src/simulator.cu:980-982).
constexpr double kTsMinEmpty = 9999999.9;
constexpr double kTsMaxEmpty = -1.0;

// Step-limit, radius and time stop rules after an accepted step; the count
// rule is applied after the removal that changes the count.
__host__ __device__ inline bool stop_reached(const Dyn &d) {
    if (d.step_limit > 0 && d.steps >= d.step_limit) return true;
    if (d.r_stop > 0 && d.max_r >= d.r_stop) return true;
    return d.time_stop &&
           (d.time_int > d.stop_int || (d.time_int == d.stop_int && d.time_frac >= d.stop_frac));
}

struct Foam {
    Consts c;
    Roles r;
    Scratch s;
    Dyn d;
    int32_t cap;         // bubble capacity
    int32_t nbr_cap;     // neighbour slots per bubble
    int32_t cell_cap;    // cell capacity
    int32_t nblk;        // bubble-space blocks
    int32_t ipt;         // bubbles per thread in bubble-space kernels
    int32_t cblk;        // cell-space blocks
    int32_t cipt;        // cells per thread in cell-space kernels
    int32_t sblk;        // neighbour-search tiles (workgroups)
    int32_t stile;       // bubbles per neighbour-search tile
    int32_t precision;
};

// launch flags
enum Flags : uint32_t {
    kGas = 1u,
    kPath = 2u,
    kCommit = 4u,
    kFlow = 8u,
};

}  // namespace cubble_amd
