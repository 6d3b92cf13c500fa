// search_kernels.hip -- searchNeighbors (reference src/simulator.cu:60-178).
//
//   k_s_cells   : wrap positions into the box on periodic axes, cell id of
//                 every bubble, cell counts (reference wrapOverPeriodicBoundaries
//                 src/kernels.cu:992-1011, cellByPosition :598-607)
//   k_s_scan1   : per-block cell-count sums
//   k_s_scan2   : cell offsets (exclusive scan), cursors, counts reset
//                 (reference: cub::DeviceScan::InclusiveSum, simulator.cu:90-91)
//   k_s_bin     : bubbles into unsorted cell segments (reference indexByCell
//                 src/kernels.cu:609-615)
//   k_s_reorder : each bubble ranks itself among its cell's members by old
//                 index (a stable counting sort, where the reference's
//                 atomic order is arbitrary) and scatters its whole state to
//                 its new slot in the alternate buffers; saved <- current
//                 (reference reorganizeByIndex :721-770)
//   k_s_search  : full neighbour list of every bubble, ascending, over the
//                 de-duplicated 3^D cell neighbourhood, one wavefront per
//                 cell (reference neighborSearch / comparePair
//                 src/kernels.cu:617-701 with getNeighborCellIndex
//                 :1116-1189); the reference's half list is the subset j > i.
//                 Counts beyond the list capacity are reported, never written
//                 out of bounds.
//   k_s_fin     : one block per foam makes the reordered buffers current
//                 (the pointer rotation of simulator.cu:96-147) and totals the
//                 pair count
#include <type_traits>

#include "dev_util.hpp"
#include "launch.hpp"

namespace cubble_amd {

// Runs of equal keys among the wave's valid lanes (bubbles are in the cell
// order of the previous search, so neighbouring lanes mostly share a cell):
// `heads` has a bit for the first lane of each run; a head's run ends at the
// next head or at the last valid lane.  One atomic per run instead of one per
// bubble.
struct WaveRuns {
    unsigned long long heads;
    int len;   // run length (valid on head lanes)
    int head;  // lane of this lane's run head
};

__device__ __forceinline__ WaveRuns wave_runs(int key, bool valid) {
    const int lane = (int)(threadIdx.x & 63);
    const int prev = __shfl_up(key, 1);
    const bool is_head = valid && (0 == lane || prev != key);
    const unsigned long long heads = __builtin_amdgcn_ballot_w64(is_head);
    const unsigned long long vmask = __builtin_amdgcn_ballot_w64(valid);
    // valid lanes are a prefix of the wave (consecutive bubbles below n)
    const int nvalid = __popcll(vmask);
    const unsigned long long above = lane < 63 ? heads & (~0ull << (lane + 1)) : 0ull;
    const int end = above ? __builtin_ctzll(above) : nvalid;
    const unsigned long long upto = heads & (lane < 63 ? (2ull << lane) - 1ull : ~0ull);
    WaveRuns r;
    r.heads = heads;
    r.len = end - lane;
    r.head = upto ? 63 - __builtin_clzll(upto) : 0;
    return r;
}

template <int D>
__device__ __forceinline__ double cells_range(const Consts &c, const Roles &R, const Scratch &S,
                                              const int cd[3], int base, int ipt, int n) {
    double rmax = 0.0;
    for (int k = 0; k < ipt; ++k) {
        if (base + k * kBlock >= n) break;  // uniform per block
        const int i = base + k * kBlock + (int)threadIdx.x;
        const bool valid = i < n;
        int cid = -1;
        if (valid) {
            double p[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < D; ++q) {
                double xv = R.x[q][i];
                if (!c.wall[q]) {
                    const int mult = xv < c.lbb[q] ? 1 : (xv > c.tfr[q] ? -1 : 0);
                    xv = xv + (c.tfr[q] - c.lbb[q]) * (double)mult;
                    R.x[q][i] = xv;
                    if (mult) R.wc[q][i] -= mult;
                }
                p[q] = xv;
            }
            int id[3] = {0, 0, 0};
#pragma unroll
            for (int q = 0; q < D; ++q) {
                const double f = floor(cd[q] * (p[q] - c.lbb[q]) / c.interval[q]);
                id[q] = f < 0.0 ? 0 : (f >= (double)cd[q] ? cd[q] - 1 : (int)f);
            }
            cid = id[2] * cd[0] * cd[1] + id[1] * cd[0] + id[0];
            S.cell_of[i] = cid;
            rmax = fmax(rmax, R.x[3][i]);
        }
        const WaveRuns wr = wave_runs(cid, valid);
        if (valid && (int)(threadIdx.x & 63) == wr.head) atomicAdd(&S.cell_count[cid], wr.len);
    }
    return rmax;
}

__global__ __launch_bounds__(kBlock) void k_s_cells(Foam *foams, const int2 *bt) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (!F->d.need_search) return;
    const Consts c = F->c;
    const Roles R = F->r;
    const Scratch S = F->s;
    const int n = F->d.count, ipt = F->ipt;
    int cd[3];
    cell_dims(c, F->d.max_r_cell, F->cell_cap, cd);
    const int base = cx.lb * kBlock * ipt;
    double rmax = 3 == c.dim ? cells_range<3>(c, R, S, cd, base, ipt, n)
                             : cells_range<2>(c, R, S, cd, base, ipt, n);
    // largest radius of the foam, for the search's reach bound: radii are
    // positive, so their bit patterns order like the values
    rmax = block_max(rmax);
    if (0 == threadIdx.x && rmax > 0.0)
        atomicMax(reinterpret_cast<unsigned long long *>(&F->d.search_max_r),
                  (unsigned long long)__double_as_longlong(rmax));
    // every block computed the same grid; no block of this kernel reads it back
    if (0 == cx.lb && 0 == threadIdx.x) {
        F->d.cell_dim[0] = cd[0];
        F->d.cell_dim[1] = cd[1];
        F->d.cell_dim[2] = cd[2];
        F->d.num_cells = cd[0] * cd[1] * cd[2];
    }
}

__global__ __launch_bounds__(kBlock) void k_s_scan1(Foam *foams, const int2 *bt) {
    const Ctx cx = cell_ctx(foams, bt);
    Foam *F = cx.F;
    if (!F->d.need_search) return;
    const Scratch S = F->s;
    const int ncell = F->d.num_cells, cipt = F->cipt;
    const int base = cx.lb * kBlock * cipt;
    int s = 0;
    for (int k = 0; k < cipt; ++k) {
        const int q = base + k * kBlock + (int)threadIdx.x;
        if (q < ncell) s += S.cell_count[q];
    }
    s = block_sum_i(s);
    if (threadIdx.x == 0) S.cpart[cx.lb] = s;
}

// Cell offsets: each block sums the block totals before it (at most a few
// hundred ints), then scans its own cells.  The block holding the last cell
// also writes start[num_cells].
__global__ __launch_bounds__(kBlock) void k_s_scan2(Foam *foams, const int2 *bt) {
    const Ctx cx = cell_ctx(foams, bt);
    Foam *F = cx.F;
    if (!F->d.need_search) return;
    const Scratch S = F->s;
    const int ncell = F->d.num_cells, cipt = F->cipt;
    const int base = cx.lb * kBlock * cipt;
    if (base >= ncell) return;  // uniform per block
    int acc = 0;
    for (int b = threadIdx.x; b < cx.lb; b += kBlock) acc += S.cpart[b];
    int off = block_sum_i(acc);
    for (int k = 0; k < cipt; ++k) {
        const int q = base + k * kBlock + (int)threadIdx.x;
        const int v = q < ncell ? S.cell_count[q] : 0;
        int tot = 0;
        const int pre = block_excl_scan(v, tot);
        if (q < ncell) {
            S.cell_start[q] = off + pre;
            S.cell_cursor[q] = off + pre;
            S.cell_count[q] = 0;
        }
        off += tot;
    }
    if (ncell - 1 < base + kBlock * cipt && threadIdx.x == 0) S.cell_start[ncell] = off;
}

__global__ __launch_bounds__(kBlock) void k_s_bin(Foam *foams, const int2 *bt) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (!F->d.need_search) return;
    const Scratch S = F->s;
    const int n = F->d.count, ipt = F->ipt;
    const int base = cx.lb * kBlock * ipt;
    for (int k = 0; k < ipt; ++k) {
        if (base + k * kBlock >= n) break;  // uniform per block
        const int i = base + k * kBlock + (int)threadIdx.x;
        const bool valid = i < n;
        const int cid = valid ? S.cell_of[i] : -1;
        // one atomic per run of equal cells; the run's lanes take consecutive
        // slots (the order inside a cell segment does not matter: the reorder
        // ranks by old index)
        const WaveRuns wr = wave_runs(cid, valid);
        const int lane = (int)(threadIdx.x & 63);
        int slot = 0;
        if (valid && lane == wr.head) slot = atomicAdd(&S.cell_cursor[cid], wr.len);
        slot = __shfl(slot, wr.head) + (lane - wr.head);
        if (valid) S.members[slot] = i;
    }
}

template <int D, typename V>
__device__ __forceinline__ void reorder_range(const Roles &R, const Scratch &S, int base, int ipt,
                                              int n) {
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        const int c = S.cell_of[i];
        const int s0 = S.cell_start[c], s1 = S.cell_start[c + 1];
        // rank among the cell's members by old index; the members are loaded
        // eight at a time (independent loads: crowded cells hold a hundred
        // bubbles and more)
        int rank = 0;
        for (int s = s0; s < s1; s += 8) {
            int mb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) mb[u] = s + u < s1 ? S.members[s + u] : i;
#pragma unroll
            for (int u = 0; u < 8; ++u) rank += mb[u] < i ? 1 : 0;
        }
        const int np = s0 + rank;
        double xs[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (2 == q && 2 == D) continue;
            const double xv = R.x[q][i];
            xs[q] = xv;
            R.xp[q][np] = xv;
            field<V>(R.saved[q])[np] = (V)xv;
            field<V>(R.vp[q])[np] = field<V>(R.v[q])[i];
            field<V>(R.vs[q])[np] = field<V>(R.vo[q])[i];
        }
        S.x4[np] = D4{xs[0], xs[1], xs[2], xs[3]};
        R.path_new[np] = R.path[i];
        field<V>(R.err_alt)[np] = field<V>(R.err)[i];
#pragma unroll
        for (int q = 0; q < D; ++q) R.wc_alt[q][np] = R.wc[q][i];
        R.idx_alt[np] = R.idx[i];
    }
}

__global__ __launch_bounds__(kBlock) void k_s_reorder(Foam *foams, const int2 *bt) {
    const Ctx cx = bubble_ctx(foams, bt);
    Foam *F = cx.F;
    if (!F->d.need_search) return;
    const Roles R = F->r;
    const Scratch S = F->s;
    const int n = F->d.count, ipt = F->ipt;
    const int base = cx.lb * kBlock * ipt;
    const bool fvel = CUBBLE_PREC_MIXED == F->precision;  // FP32 velocities
    if (3 == F->c.dim) {
        if (fvel) reorder_range<3, float>(R, S, base, ipt, n);
        else reorder_range<3, double>(R, S, base, ipt, n);
    } else {
        if (fvel) reorder_range<2, float>(R, S, base, ipt, n);
        else reorder_range<2, double>(R, S, base, ipt, n);
    }
}

// After the reorder the new state sits in the alternate buffers; k_s_fin
// makes them current.
__device__ void swap_reordered_roles(Roles &W) {
    for (int q = 0; q < 4; ++q) {
        double *t = W.x[q];
        W.x[q] = W.xp[q];
        W.xp[q] = t;
        void *u = W.v[q];
        W.v[q] = W.vp[q];
        W.vp[q] = u;
        u = W.vo[q];
        W.vo[q] = W.vs[q];
        W.vs[q] = u;
    }
    double *t = W.path;
    W.path = W.path_new;
    W.path_new = t;
    void *e = W.err;
    W.err = W.err_alt;
    W.err_alt = e;
    for (int q = 0; q < 3; ++q) {
        int *u = W.wc[q];
        W.wc[q] = W.wc_alt[q];
        W.wc_alt[q] = u;
    }
    int *u = W.idx;
    W.idx = W.idx_alt;
    W.idx_alt = u;
}

// Sorted, de-duplicated neighbour coordinates of cell coordinate `x` along
// one axis: periodic axes wrap, wall axes drop out-of-range cells.  With
// fewer than 3 cells the wrapped offsets coincide and appear once.  Held in
// three named values, not an array, so no kernel indexes private memory.
struct AxisCells {
    int v0, v1, v2, n;
    __device__ __forceinline__ int at(int k) const { return 0 == k ? v0 : 1 == k ? v1 : v2; }
};

__device__ __forceinline__ AxisCells axis_cells(int x, int n, bool wall) {
    AxisCells r;
    if (wall) {
        const bool lo = x - 1 >= 0, hi = x + 1 < n;
        r.v0 = lo ? x - 1 : x;
        r.v1 = lo ? x : x + 1;
        r.v2 = x + 1;
        r.n = 1 + (lo ? 1 : 0) + (hi ? 1 : 0);
        return r;
    }
    int v0 = (x - 1 + n) % n, v1 = x, v2 = (x + 1) % n, t;
    if (v0 > v1) { t = v0; v0 = v1; v1 = t; }
    if (v1 > v2) { t = v1; v1 = v2; v2 = t; }
    if (v0 > v1) { t = v0; v0 = v1; v1 = t; }
    // drop duplicates, keeping order
    const bool d1 = v1 == v0, d2 = v2 == v1;
    r.v0 = v0;
    r.v1 = d1 ? v2 : v1;
    r.v2 = v2;
    r.n = 3 - (d1 ? 1 : 0) - (d2 ? 1 : 0);
    return r;
}

// The x-cells of one row as at most two runs of consecutive cells,
// [a0, b0] and (if two) [a1, b1], in ascending order.
struct XRuns {
    int a0, b0, a1, b1;
    bool two;
};

__device__ __forceinline__ XRuns x_runs(const AxisCells &X) {
    int f0 = 0;
    if (X.n > 1 && X.v1 == X.v0 + 1) f0 = 1;
    if (1 == f0 && X.n > 2 && X.v2 == X.v1 + 1) f0 = 2;
    XRuns r;
    r.a0 = X.v0;
    r.b0 = X.at(f0);
    r.two = f0 + 1 < X.n;
    r.a1 = r.two ? X.at(f0 + 1) : 0;
    r.b1 = r.two ? X.at(X.n - 1) : -1;
    return r;
}

// The x-cells [x0 - 1, x0 + k] around k consecutive own cells x0 .. x0 + k - 1
// of one row, as at most two runs in ascending order, each with its image
// code as `side` gives it for a neighbour cell (0; 1: the run lies past the
// high end, its candidates take d - L; 2: past the low end, d + L).  On a
// periodic axis k + 4 <= n, so the window wraps at one end at most and a
// cell appears once.  A pair of cells more than one apart in the window is
// at least a cell width apart under either image, so it never holds a pair
// within reach.
struct XWindow {
    int a0, b0, f0, a1, b1, f1;
    bool two;
};

__device__ __forceinline__ XWindow x_window(int x0, int k, int n, bool wall) {
    XWindow r{x0 - 1, x0 + k, 0, 0, -1, 0, false};
    if (wall) {
        r.a0 = max(r.a0, 0);
        r.b0 = min(r.b0, n - 1);
    } else if (x0 - 1 < 0) {  // [0, x0 + k], then cell n - 1 past the low end
        r = XWindow{0, x0 + k, 0, n - 1, n - 1, 2, true};
    } else if (x0 + k >= n) {  // cells past the high end first, then [x0 - 1, n - 1]
        r = XWindow{0, x0 + k - n, 1, x0 - 1, n - 1, 0, true};
    }
    return r;
}

// Wave-uniform append of the ballot `m` (non-zero) of candidates j to own
// bubble i's list.  Lane `ii` of the group (ii is wave-uniform) keeps the
// count: it is read with readlane and updated on the scalar unit.  A hit
// lane's slot is the count plus the hits in the lanes below it (mbcnt); the
// hit itself is re-derived from `m`, so the comparison result need not stay
// live across the branch.
__device__ __forceinline__ void append_hits(unsigned long long m, int ii, int i, int j,
                                            int *__restrict__ nbr_, unsigned cap, int nbr_cap,
                                            int &cnt) {
    const int base = __builtin_amdgcn_readlane(cnt, ii);
    const int upd = base + __popcll(m);  // scalar
    // the hit lanes become the exec mask directly (no per-lane bit test)
    if (__builtin_amdgcn_inverse_ballot_w64(m)) {
        const int below = (int)__builtin_amdgcn_mbcnt_hi(
            (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        const int pos = base + below;
        auto nbr = (__attribute__((address_space(1))) int *)(nbr_ + i);
        // slots past the capacity are counted, not written; when all fit (a
        // scalar test) the bound is one no slot reaches
        const int lim = upd <= nbr_cap ? 0x7fffffff : nbr_cap;
        if (pos < lim) nbr[(uint64_t)(unsigned)pos * cap] = j;
    }
    cnt = (int)(threadIdx.x & 63) == ii ? upd : cnt;  // lane ii keeps the count
}

// The list length goes to nn and, as the degree the imposed flow divides by,
// to deg (which a removal moves but does not reduce).
__device__ __forceinline__ void finish_group(int cb, int nb, int cnt, int nbr_cap,
                                             int *__restrict__ nn, int *__restrict__ deg,
                                             int &sum_nn, int &max_nn, int &ovf) {
    if ((int)(threadIdx.x & 63) < nb) {
        const int v = cnt < nbr_cap ? cnt : nbr_cap;
        nn[cb + (int)(threadIdx.x & 63)] = v;
        deg[cb + (int)(threadIdx.x & 63)] = v;
        sum_nn += cnt;
        max_nn = max(max_nn, cnt);
        if (cnt > nbr_cap) ovf = 1;
    }
}

// Neighbour lists, one wavefront per cell, any box (the general path).
// The linear cell id z*cx*cy + y*cx + x grows lexicographically in
// (z, y, x), so walking the sorted per-axis cell lists in that nesting
// visits cells -- and, since bubbles are stored in cell order, candidates --
// in ascending index order; consecutive x-cells of one row form one
// contiguous run of bubbles.  The 64 lanes take 64 consecutive candidates of
// a run (one coalesced 32-byte record each); the cell's own bubbles are
// wave-uniform scalar loads; hits are found with a wave ballot and appended
// to that bubble's list in lane order, which keeps every list ascending.
// Lists longer than the capacity are counted, never written out of bounds.
// `cell` must be wave-uniform.
template <int D>
__device__ __forceinline__ void search_cell_general(const Consts &c, const D4 *__restrict__ x4_,
                                                    const int *__restrict__ start_,
                                                    int *__restrict__ nbr, int *__restrict__ nn,
                                                    int *__restrict__ deg,
                                                    const int cd[3], size_t cap, int nbr_cap,
                                                    int cell, int i0, int i1, int &sum_nn,
                                                    int &max_nn, int &ovf) {
    const auto x4 = gmem(x4_);
    const auto x4s = cmem(x4_);
    const auto start = cmem(start_);
    const int lane = (int)(threadIdx.x & 63);
    cell = uniform_i(cell);
    i0 = uniform_i(i0);
    i1 = uniform_i(i1);
    if (i0 >= i1) return;
    const int ccx = cell % cd[0];
    const int ccy = (cell / cd[0]) % cd[1];
    const int ccz = cell / (cd[0] * cd[1]);
    const AxisCells X = axis_cells(ccx, cd[0], c.wall[0] != 0);
    const AxisCells Y = axis_cells(ccy, cd[1], c.wall[1] != 0);
    const AxisCells Z = 3 == D ? axis_cells(ccz, cd[2], c.wall[2] != 0) : AxisCells{0, 0, 0, 1};
    const XRuns xr = x_runs(X);
    const double skin = c.skin;
    for (int cb = i0; cb < i1; cb += 64) {
        const int nb = uniform_i(min(64, i1 - cb));
        int cnt = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int b = 0; b < 3; ++b) {
#pragma unroll
                for (int part = 0; part < 2; ++part) {
                    if (a >= Z.n || b >= Y.n || (1 == part && !xr.two)) continue;
                    const int row = Z.at(a) * cd[0] * cd[1] + Y.at(b) * cd[0];
                    const int j0 = start[row + (0 == part ? xr.a0 : xr.a1)];
                    const int j1 = start[row + (0 == part ? xr.b0 : xr.b1) + 1];
                    for (int jb = j0; jb < j1; jb += 64) {
                        const int j = jb + lane;
                        const bool valid = j < j1;
                        const unsigned long long vmask = __builtin_amdgcn_ballot_w64(valid);
                        const int jl = valid ? j : j0;
                        const D4 pj = D4{x4[jl].x, x4[jl].y, 3 == D ? x4[jl].z : 0.0, x4[jl].r};
                        for (int ii = 0; ii < nb; ++ii) {
                            const int i = cb + ii;
                            const double px = x4s[i].x, py = x4s[i].y, pr = x4s[i].r;
                            const double pz = 3 == D ? x4s[i].z : 0.0;
                            const double maxd = pr + pj.r + skin;
                            double d[3];
                            wrapped_diff<D>(c, px, py, pz, pj.x, pj.y, pj.z, d);
                            // one ballot per comparison: each is the comparison's own
                            // lane mask, with no boolean materialised in a register
                            const unsigned long long m =
                                vmask & __builtin_amdgcn_ballot_w64(j != i) &
                                __builtin_amdgcn_ballot_w64(sq_len(d) < maxd * maxd);
                            if (m) append_hits(m, ii, i, j, nbr, (unsigned)cap, nbr_cap, cnt);
                        }
                    }
                }
            }
        }
        finish_group(cb, nb, cnt, nbr_cap, nn, deg, sum_nn, max_nn, ovf);
    }
}

constexpr int kMaxRuns = 18;  // 3 z-rows x 3 y-rows x 2 x-runs
constexpr int kOwnUnit = 32;  // own bubbles per search work unit
constexpr int kWin2 = 4;      // 2D: own cells of one row per work unit

// Neighbour lists of one cell when every periodic axis has at least 5 cells
// (the common case).  The candidate runs of the cell (at most two per row:
// the x-cells of a row split only where the row wraps) are concatenated, in
// ascending bubble order, into one sequence that the 64 lanes walk 64 at a
// time, so partly filled chunks occur once per cell instead of once per run.
// With 5 or more cells per axis the minimum image is decided by the cells
// alone: a candidate in a neighbouring cell that does not cross the boundary
// is within 2/5 of the box and needs no wrap (|d| < L - |d|), one across the
// boundary is at least 3/5 away and always takes the image, whose sign the
// cells fix: d - L for a candidate across the low end (d > 0), d + L across
// the high end.  Rounding is symmetric, so fl(d - L) = -fl(L - |d|) and
// fl(d + L) = fl(L - |d|): the values of the general minimum-image
// difference, without its test.  Same order, same lists, same rounding as
// search_cell_general.
//
// Pruning by reach (3D): an own bubble i cannot hit a candidate in the
// neighbouring row of cells below it in z (or y) when its distance to that
// face of its cell is at least r_i + rmax + skin, rmax the largest radius of
// the foam (all candidates lie beyond the face).  Each lane holds, for one
// own bubble, the rows of the 3 x 3 (z, y) stencil it can reach; a chunk of
// candidates is tested only against the own bubbles that reach a row the
// chunk touches.  The reach is widened by 1e-9 (relative, and of the box)
// against rounding, so no hit is ever skipped: same lists.  rmax <= 0 (not
// known) disables the pruning.
// `cell` must be wave-uniform.
template <int D>
__device__ __forceinline__ void search_cell(const Consts &c, const D4 *__restrict__ x4_,
                                            const int *__restrict__ start_,
                                            int *__restrict__ nbr, int *__restrict__ nn,
                                            int *__restrict__ deg,
                                            const int cd[3], size_t cap, int nbr_cap, int cell,
                                            int kx, int i0, int i1, double rmax, int &sum_nn,
                                            int &max_nn, int &ovf) {
    const auto x4 = gmem(x4_);
    const auto x4s = cmem(x4_);
    const auto start = cmem(start_);
    const int lane = (int)(threadIdx.x & 63);
    cell = uniform_i(cell);
    kx = uniform_i(kx);
    i0 = uniform_i(i0);
    i1 = uniform_i(i1);
    if (i0 >= i1) return;
    const int ccx = cell % cd[0];
    const int ccy = (cell / cd[0]) % cd[1];
    const int ccz = cell / (cd[0] * cd[1]);
    const AxisCells X = axis_cells(ccx, cd[0], c.wall[0] != 0);
    const AxisCells Y = axis_cells(ccy, cd[1], c.wall[1] != 0);
    const AxisCells Z = 3 == D ? axis_cells(ccz, cd[2], c.wall[2] != 0) : AxisCells{0, 0, 0, 1};
    const XRuns xr = x_runs(X);
    // 0: neighbour cell a of own cell b does not cross the boundary; 1: it
    // lies across the low end (d > 0, image d - L); 2: across the high end
    // (d < 0, image d + L)
    auto side = [](int a, int b) { return b - a > 1 ? 1 : (a - b > 1 ? 2 : 0); };
    const int fx0 = side(xr.a0, ccx), fx1 = side(xr.a1, ccx);
    // 2D foams have one z-row: 6 runs
    constexpr int kRuns = 3 == D ? kMaxRuns : kMaxRuns / 3;
    // 2D: the runs' starts, offsets and image codes in scalar registers;
    // 3D: one run per lane (lane s < kRuns holds run s, the lanes above it
    // empty runs at the end), read by the chunks with readlane
    int off[kRuns], pre[kRuns], fl[kRuns];
    int lane_pre = 0, lane_off = 0, lane_fl = 0;
    int total = 0;
    if (3 == D) {
        const int r = lane >> 1, part = lane & 1;
        const int a = r / 3, b = r % 3;
        const bool ok = lane < kRuns && a < Z.n && b < Y.n && (0 == part || xr.two);
        const int zc = Z.at(a), yc = Y.at(b);
        const int row = zc * cd[0] * cd[1] + yc * cd[0];
        const auto vstart = gmem(start_);
        const int js = ok ? vstart[row + (0 == part ? xr.a0 : xr.a1)] : 0;
        const int je = ok ? vstart[row + (0 == part ? xr.b0 : xr.b1) + 1] : 0;
        const int len = je - js;
        // run starts: the exclusive prefix of the lengths over the lanes
        int incl = len;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        lane_pre = incl - len;
        total = __builtin_amdgcn_readlane(incl, kRuns - 1);
        lane_off = js - lane_pre;  // candidate p of the run is bubble p + off
        lane_fl = (side(yc, ccy) << 2) | (side(zc, ccz) << 4) | (0 == part ? fx0 : fx1);
    }
    // 2D: the x-window of the unit's kx own cells of one row (kx = 1: the
    // cell and its two neighbours, as in 3D)
    const XWindow xw = x_window(ccx, kx, cd[0], c.wall[0] != 0);
#pragma unroll
    for (int a = 0; a < (3 == D ? 0 : 1); ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const bool row_ok = a < Z.n && b < Y.n;
            const int zc = Z.at(a), yc = Y.at(b);
            const int row = zc * cd[0] * cd[1] + yc * cd[0];
            const int fyz = (side(yc, ccy) << 2) | (3 == D ? side(zc, ccz) << 4 : 0);
#pragma unroll
            for (int part = 0; part < 2; ++part) {
                const int s = (a * 3 + b) * 2 + part;
                const bool ok = row_ok && (0 == part || xw.two);
                const int js = ok ? start[row + (0 == part ? xw.a0 : xw.a1)] : 0;
                const int je = ok ? start[row + (0 == part ? xw.b0 : xw.b1) + 1] : 0;
                // candidate p of the run is bubble p + off
                off[s] = js - total;
                pre[s] = total;
                fl[s] = fyz | (0 == part ? xw.f0 : xw.f1);
                total += je - js;
            }
        }
    }
    const double skin = c.skin;
    const double L[3] = {c.interval[0], c.interval[1], c.interval[2]};
    // rows of the (z, y) stencil: row r = a * 3 + b holds runs 2r, 2r + 1;
    // its offsets from the own cell in z and y (-1, 0, +1)
    constexpr int kRows = kRuns / 2;
    auto offset = [](int v, int cc, int n) {
        int o = v - cc;
        if (o > 1) o -= n;
        if (o < -1) o += n;
        return o;
    };
    // own cell faces in y and z; widened reach against rounding
    const double ey = L[1] / cd[1], ez = L[2] / cd[2];
    const double ylo = c.lbb[1] + ccy * ey, yhi = ylo + ey;
    const double zlo = c.lbb[2] + ccz * ez, zhi = zlo + ez;
    const double slack = 1e-9 * fmax(L[0], fmax(L[1], L[2]));
    // 2D cells hold about three bubbles, so a cell's candidates are one chunk
    // touching all three rows, which no own bubble can skip: pruning in 3D only
    const bool prune = 3 == D && rmax > 0.0;
    for (int cb = i0; cb < i1; cb += 64) {
        const int nb = uniform_i(min(64, i1 - cb));
        int cnt = 0;
        // rows each own bubble (one per lane) can reach
        int reach_rows = (1 << kRows) - 1;
        if (prune) {
            const int o = cb + (lane < nb ? lane : nb - 1);
            const double py = x4[o].y, pz = 3 == D ? x4[o].z : 0.0;
            const double reach = (x4[o].r + rmax + skin) * (1.0 + 1e-9) + slack;
            const bool ry_lo = py - ylo < reach, ry_hi = yhi - py < reach;
            const bool rz_lo = 3 == D && pz - zlo < reach, rz_hi = 3 == D && zhi - pz < reach;
            reach_rows = 0;
#pragma unroll
            for (int a = 0; a < (3 == D ? 3 : 1); ++a) {
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const int oz = 3 == D && a < Z.n ? offset(Z.at(a), ccz, cd[2]) : 0;
                    const int oy = b < Y.n ? offset(Y.at(b), ccy, cd[1]) : 0;
                    const bool okz = 0 == oz || (oz < 0 ? rz_lo : rz_hi);
                    const bool oky = 0 == oy || (oy < 0 ? ry_lo : ry_hi);
                    if (okz && oky) reach_rows |= 1 << (a * 3 + b);
                }
            }
        }
        const unsigned long long own_all = nb >= 64 ? ~0ull : (1ull << nb) - 1ull;
        const auto own = x4s + cb;  // the group's own records (scalar loads)
        int *const nbr_own = nbr + cb;  // their list columns
        int s_cur = 0;  // 3D: the last non-empty run the previous chunk reached
        for (int p0 = 0; p0 < total; p0 += 64) {
            const int p = p0 + lane;
            // the rows this chunk of candidates touches, and the run of
            // candidate p: the last non-empty one starting at or before p
            int rows = 0, o = 0, f = 0;
            if (3 == D) {
                // only the runs that start before the chunk's end, from the
                // one holding p0 on (scalar loop, readlane of the tables)
                int sr = s_cur;
                int rs = __builtin_amdgcn_readlane(lane_pre, sr);
                while (sr < kRuns && rs < p0 + 64) {
                    const int re = __builtin_amdgcn_readlane(lane_pre, sr + 1);
                    if (re > rs && re > p0) {
                        const int os = __builtin_amdgcn_readlane(lane_off, sr);
                        const int fs = __builtin_amdgcn_readlane(lane_fl, sr);
                        const bool in = p >= rs;
                        o = in ? os : o;
                        f = in ? fs : f;
                        rows |= 1 << (sr >> 1);
                        s_cur = sr;
                    }
                    ++sr;
                    rs = re;
                }
            } else {
                rows = (1 << kRows) - 1;  // no pruning: every own bubble takes every chunk
                o = off[0];
                f = fl[0];
#pragma unroll
                for (int s = 1; s < kRuns; ++s) {
                    const bool in = p >= pre[s];
                    o = in ? off[s] : o;
                    f = in ? fl[s] : f;
                }
            }
            const unsigned long long own_mask =
                own_all & __builtin_amdgcn_ballot_w64((reach_rows & rows) != 0);
            if (!own_mask) continue;
            const bool valid = p < total;
            const int j = valid ? p + o : -1;
            const int jl = valid ? j : i0;
            // a lane past the end holds a candidate at infinity: its
            // distance test is false, so no validity mask enters the test
            const D4 pj = D4{valid ? x4[jl].x : __builtin_inf(), x4[jl].y,
                             3 == D ? x4[jl].z : 0.0, x4[jl].r};
            const bool anyflip = __builtin_amdgcn_ballot_w64(valid && f != 0) != 0ull;
            // the self test only where the chunk holds some of the own bubbles
            const bool anyself =
                __builtin_amdgcn_ballot_w64(valid && j >= cb && j < cb + nb) != 0ull;
            // image shift of this lane's candidate per axis: 0, +L or -L
            double sh[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int fk = (f >> (2 * k)) & 3;
                sh[k] = 1 == fk ? L[k] : (2 == fk ? -L[k] : 0.0);
            }
            // the own-bubble loop in two versions, so that chunks with no
            // boundary crossing carry no image arithmetic
            // hit mask of own bubble i against the chunk (one ballot per
            // comparison: each is the comparison's own lane mask, with no
            // boolean materialised in a register)
            auto test = [&](auto flip, auto self, unsigned ia) {
                // unsigned index: no sign extension in the scalar address
                const double px = own[ia].x, py = own[ia].y, pr = own[ia].r;
                const double pz = 3 == D ? own[ia].z : 0.0;
                double d[3] = {px - pj.x, py - pj.y, 3 == D ? pz - pj.z : 0.0};
                if (decltype(flip)::value) {
                    // d - (+-L) rounds exactly as -+(L - |d|); d - 0 is d
#pragma unroll
                    for (int k = 0; k < D; ++k) d[k] = d[k] - sh[k];
                }
                const double maxd = pr + pj.r + skin;
                const unsigned long long hit =
                    __builtin_amdgcn_ballot_w64(sq_len(d) < maxd * maxd);
                if (decltype(self)::value)
                    return hit & __builtin_amdgcn_ballot_w64(j != cb + (int)ia);
                return hit;
            };
            // two own bubbles per iteration: half the loop and load overhead on
            // the scalar unit, two independent FP64 chains in flight
            auto own_loop = [&](auto flip, auto self) {
                unsigned long long todo = own_mask;
                while (todo) {
                    const int ia = __builtin_ctzll(todo);
                    todo &= todo - 1ull;
                    if (todo) {
                        const int ib = __builtin_ctzll(todo);
                        todo &= todo - 1ull;
                        const unsigned long long ma = test(flip, self, (unsigned)ia);
                        const unsigned long long mb = test(flip, self, (unsigned)ib);
                        if (ma) append_hits(ma, ia, ia, j, nbr_own, (unsigned)cap, nbr_cap, cnt);
                        if (mb) append_hits(mb, ib, ib, j, nbr_own, (unsigned)cap, nbr_cap, cnt);
                    } else {
                        const unsigned long long ma = test(flip, self, (unsigned)ia);
                        if (ma) append_hits(ma, ia, ia, j, nbr_own, (unsigned)cap, nbr_cap, cnt);
                    }
                }
            };
            if (anyflip) {
                if (anyself) own_loop(std::true_type{}, std::true_type{});
                else own_loop(std::true_type{}, std::false_type{});
            } else {
                if (anyself) own_loop(std::false_type{}, std::true_type{});
                else own_loop(std::false_type{}, std::false_type{});
            }
        }
        finish_group(cb, nb, cnt, nbr_cap, nn, deg, sum_nn, max_nn, ovf);
    }
}

// Search grid: tile t of a foam searches for the own bubbles [t B, t B + B),
// B = stile bubbles, whatever cells hold them (a cell cut by a tile
// boundary is walked by both tiles, each for its own bubbles).  Tiles sized in bubbles keep the grid
// proportional to the bubbles of a foam whatever its cell occupancy (2D
// cells hold a few bubbles, crowded 3D ones a hundred and more).
// Eight waves per SIMD: left to itself the compiler keeps more scalar
// registers (the candidate runs, the own records) than allow that, and the
// kernel ran at seven (4-6% slower in the batch's searches).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_s_search(
    Foam *foams, const int2 *bt) {
    const Ctx cx = cell_ctx(foams, bt);
    Foam *F = cx.F;
    if (!F->d.need_search) return;
    const Consts c = F->c;
    const Scratch S = F->s;
    const int cd[3] = {F->d.cell_dim[0], F->d.cell_dim[1], F->d.cell_dim[2]};
    const double rmax = F->d.search_max_r;  // 0 when not computed (list rebuild)
    // tile t: own bubbles [t B, t B + B) (B = stile), whatever cells they lie
    // in; the cells holding them are [base, end)
    const int b0 = cx.lb * F->stile;
    const int b1 = min(b0 + F->stile, F->d.count);
    // the cells holding them: position q of the cell order lies in the
    // segment of the cell of members[q] (any member of that segment), so the
    // first and last own bubbles name the range (scalar loads)
    int base = 0, end = 0;
    if (b0 < b1) {
        const auto mem = cmem(S.members);
        const auto cof = cmem(S.cell_of);
        base = cof[mem[b0]];
        end = cof[mem[b1 - 1]] + 1;
    }
    const size_t cap = (size_t)F->cap;
    int sum_nn = 0, max_nn = 0, ovf = 0;
    // the fast path needs 5 or more cells along every periodic axis
    bool fast = true;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (k < c.dim && !c.wall[k] && cd[k] < 5) fast = false;
    // work units: up to kOwnUnit consecutive own bubbles of one cell (within
    // the tile), dealt round-robin to the waves, so a crowded cell (a hundred
    // bubbles and more once the radii spread) is shared by the waves of
    // several workgroups instead of serialising one wave; each own bubble's
    // list is still built by one wave, candidates in ascending order
    //
    // 2D cells hold about three bubbles, so a unit there takes up to kWin2
    // consecutive cells of one row: its own bubbles share one candidate
    // window (the rows above and below, the cells one past either end), so
    // the unit's set-up and candidate chunk serve about kWin2 times as many
    // own bubbles; a candidate more than one cell from an own bubble's cell
    // is out of its reach and never listed.
    const int w = (int)(threadIdx.x >> 6);
    const auto cstart = cmem(S.cell_start);
    int win2 = 1;
    if (2 == c.dim && fast) win2 = c.wall[0] ? kWin2 : min(kWin2, cd[0] - 4);
    int uid = 0;
    for (int cell = base; cell < end;) {
        // own cells [cell, cell + kx): one row, within the tile's cells
        const int row_end = (cell / cd[0] + 1) * cd[0];
        const int kx = max(1, min(win2, min(end, row_end) - cell));
        const int s0 = max(cstart[cell], b0), s1 = min(cstart[cell + kx], b1);
        for (int o0 = s0; o0 < s1; o0 += kOwnUnit, ++uid) {
            if ((uid & (kWaves - 1)) != w) continue;
            const int o1 = min(o0 + kOwnUnit, s1);
            if (3 == c.dim) {
                if (fast)
                    search_cell<3>(c, S.x4, S.cell_start, S.nbr, S.nn, S.deg, cd, cap,
                                   F->nbr_cap, cell, 1, o0, o1, rmax, sum_nn, max_nn, ovf);
                else
                    search_cell_general<3>(c, S.x4, S.cell_start, S.nbr, S.nn, S.deg, cd, cap,
                                           F->nbr_cap, cell, o0, o1, sum_nn, max_nn, ovf);
            } else {
                if (fast)
                    search_cell<2>(c, S.x4, S.cell_start, S.nbr, S.nn, S.deg, cd, cap,
                                   F->nbr_cap, cell, kx, o0, o1, rmax, sum_nn, max_nn, ovf);
                else
                    search_cell_general<2>(c, S.x4, S.cell_start, S.nbr, S.nn, S.deg, cd, cap,
                                           F->nbr_cap, cell, o0, o1, sum_nn, max_nn, ovf);
            }
        }
        cell += kx;
    }
    sum_nn = block_sum_i(sum_nn);
    max_nn = block_max_i(max_nn);
    ovf = block_max_i(ovf);
    if (threadIdx.x == 0) {
        S.cpart_i[cx.lb * kPartI + 1] = sum_nn;
        S.cpart_i[cx.lb * kPartI + 2] = max_nn;
        S.cpart_i[cx.lb * kPartI + 3] = ovf;
    }
}

// One block per foam: make the reordered buffers current (swap = 1), pair
// count, list-capacity check, counters.
__global__ __launch_bounds__(kBlock) void k_s_fin(Foam *foams, int swap) {
    Foam *F = foams + blockIdx.x;
    if (!F->d.need_search) return;
    const Scratch S = F->s;
    const int nb = F->sblk;
    int s = 0, mx = 0, of = 0;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        s += S.cpart_i[b * kPartI + 1];
        mx = max(mx, S.cpart_i[b * kPartI + 2]);
        of = max(of, S.cpart_i[b * kPartI + 3]);
    }
    const int tot = block_sum_i(s);
    mx = block_max_i(mx);
    of = block_max_i(of);
    if (threadIdx.x == 0) {
        if (swap) swap_reordered_roles(F->r);
        F->d.num_pairs = tot / 2;
        F->d.max_nn = mx;
        F->d.searches += 1;
        if (of) {
            F->d.overflow = 1;
            F->d.halt = 1;
        }
        F->d.need_search = 0;
        F->d.search_max_r = 0.0;  // the next search recomputes it
    }
}

__global__ void k_set_flags(Foam *foams, int nf, int ns, int nr, int clear_halt) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    if (ns >= 0) foams[f].d.need_search = ns;
    if (nr >= 0) foams[f].d.need_remove = nr;
    if (clear_halt) {
        foams[f].d.halt = 0;
        foams[f].d.overflow = 0;
    }
}

// add > 0: every foam halts after `add` more accepted steps; add == 0:
// no limit.  Clears halt / overflow.
__global__ void k_set_limits(Foam *foams, int nf, long long add) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    Dyn &d = foams[f].d;
    d.step_limit = add > 0 ? d.steps + add : 0;
    d.halt = 0;
    d.overflow = 0;
}

void launch_set_limits(const Grid &g, long long add) {
    const int nb = (g.nfoams + 63) / 64;
    hipLaunchKernelGGL(k_set_limits, dim3(nb), dim3(64), 0, g.stream, g.foams, g.nfoams, add);
}

void launch_search(const Grid &g) {
    hipLaunchKernelGGL(k_s_cells, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams, g.bt_bub);
    hipLaunchKernelGGL(k_s_scan1, dim3(g.nb_cell), dim3(kBlock), 0, g.stream, g.foams, g.bt_cell);
    hipLaunchKernelGGL(k_s_scan2, dim3(g.nb_cell), dim3(kBlock), 0, g.stream, g.foams, g.bt_cell);
    hipLaunchKernelGGL(k_s_bin, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams, g.bt_bub);
    hipLaunchKernelGGL(k_s_reorder, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams, g.bt_bub);
    hipLaunchKernelGGL(k_s_search, dim3(g.nb_scell), dim3(kBlock), 0, g.stream, g.foams,
                       g.bt_scell);
    hipLaunchKernelGGL(k_s_fin, dim3(g.nfoams), dim3(kBlock), 0, g.stream, g.foams, 1);
}

void launch_search_lists(const Grid &g) {
    hipLaunchKernelGGL(k_s_search, dim3(g.nb_scell), dim3(kBlock), 0, g.stream, g.foams,
                       g.bt_scell);
    hipLaunchKernelGGL(k_s_fin, dim3(g.nfoams), dim3(kBlock), 0, g.stream, g.foams, 0);
}

void launch_search_part(const Grid &g, int part) {
    if (kPartReorder == part)
        hipLaunchKernelGGL(k_s_reorder, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams,
                           g.bt_bub);
    else  // the list build is the list kernel alone
        hipLaunchKernelGGL(k_s_search, dim3(g.nb_scell), dim3(kBlock), 0, g.stream, g.foams,
                           g.bt_scell);
}

void launch_set_flags(const Grid &g, int need_search, int need_remove, int clear_halt) {
    const int nb = (g.nfoams + 63) / 64;
    hipLaunchKernelGGL(k_set_flags, dim3(nb), dim3(64), 0, g.stream, g.foams, g.nfoams, need_search,
                       need_remove, clear_halt);
}

}  // namespace cubble_amd
