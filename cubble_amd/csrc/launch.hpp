// launch.hpp -- host-side launch wrappers of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "foam.hpp"

namespace cubble_amd {

// A launch covers one or more foams.  bt_bub[b] = (foam, local block) for
// the bubble-space grid, bt_cell[b] likewise for the cell-space grid.
struct Grid {
    Foam *foams;
    const int2 *bt_bub;
    int nb_bub;
    const int2 *bt_cell;
    int nb_cell;
    const int2 *bt_scell;  // neighbour-search grid (Foam::stile bubbles per tile)
    int nb_scell;
    int nfoams;
    hipStream_t stream;
    bool walls;  // some foam of the launch has a wall
};

void launch_predict(const Grid &g, uint32_t flags);
void launch_pair(const Grid &g, uint32_t flags);
// The mixed pair kernel, then (fix) the second pass over the bubbles that met
// a pair needing FP64 values.
void launch_pair_mixed(const Grid &g, uint32_t flags, bool fix = true);
// S_A, S_Ar, S_OA, S_OAr from the predictor and pair partials (one block per foam)
void launch_sums(const Grid &g);
// corrector (mixed: the mixed-precision variant, for engines of that
// precision), then (control) the dt controller / commit (one block per foam)
void launch_correct(const Grid &g, uint32_t flags, bool mixed, bool control = true);

// removeBubbles: emit the sorted deletion list and plan the moves, then move
// the tail survivors into the holes, filter / remap the neighbour lists and
// apply the volume fix.  Runs only for foams whose need_remove flag is set.
void launch_remove(const Grid &g);

// searchNeighbors: wrap + cell ids + counts, two-pass scan, binning,
// rank-and-scatter reorder, neighbour lists.  Runs only for foams whose
// need_search flag is set.
void launch_search(const Grid &g);
// Only the neighbour-list kernels (after the list capacity was raised).
void launch_search_lists(const Grid &g);
// One part of the search, for timing: the list build (the list kernel alone)
// or the reorder.
enum SearchPart { kPartLists = 0, kPartListKernel = 1, kPartReorder = 2 };
void launch_search_part(const Grid &g, int part);

// Set / clear the phase flags of every foam in the launch.
void launch_set_flags(const Grid &g, int need_search, int need_remove, int clear_halt);
// add > 0: every foam halts after `add` more accepted steps; 0: no limit.
void launch_set_limits(const Grid &g, long long add);

// The AB2 velocity bootstrap of init: pair velocities (no gas, flow or
// walls) at the positions moved by one Euler step of the pair velocities
// become the old velocities.  Needs zero velocities (a new foam).
void launch_bootstrap(const Grid &g, bool mixed);

// Positions of one foam moved from its box (h.c) to the box of new_c, each
// coordinate keeping its relative place; the constants are the caller's.
void launch_rescale(const Foam &h, Foam *dev, const Consts &new_c, hipStream_t stream);

// One foam's observables (h: a current host copy of its header, dev: its
// device header): out[0..7] = total energy, total volume, then the sums of
// r, path, dxdt, dydt, dzdt, drdt; energy[i] (if not null) = bubble i's
// energy.  Uses the foam's block partials.
void launch_observe(const Foam &h, Foam *dev, double *out, double *energy, hipStream_t stream);

}  // namespace cubble_amd
