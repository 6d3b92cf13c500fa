// pairs.hip -- the reference's searchNeighbors output (src/simulator.cu:60-178)
// as a function of caller-owned device arrays: cubble_find_pairs.
//
// The reference fills Pairs {i, j} (src/data_definitions.h:157-178) with every
// pair i < j whose minimum-image distance is below r_i + r_j + skin
// (comparePair, src/kernels.cu:617-634), grouped by first index (sortPairs,
// :703-711).  Here the caller hands over x, y, z, r (device, FP64) and gets
// pair_i / pair_j in that order, each group also ascending in j, plus the
// pair count in device memory: nothing is read back to the host and the
// host never waits.  The bubbles are not moved: positions are wrapped into the box on the fly (the
// reference wraps them in place first, wrapOverPeriodicBoundaries :992-1011).
//
//   k_p_init  : clears the radius maximum (no memset: the calls may be captured)
//   k_p_maxr  : largest radius, for the cell size 2 (max r + skin)
//               (simulator.cu:69-75)
//   k_p_grid  : cell grid (every block derives the same one), cell counts zeroed
//   k_p_cells : cell of every bubble, cell counts (cellByPosition :598-607)
//   k_p_scan* : exclusive scans (cell starts; per-bubble pair offsets)
//   k_p_bin   : bubbles into cell segments (indexByCell :609-615)
//   k_p_walk  : per bubble, the candidates j > i of the de-duplicated 3^D
//               stencil: first counted, then written and sorted
//   k_p_total : the pair count
//
// Scratch (cell arrays, per-bubble counts and offsets, scan partials) is
// sized from the bubble count alone: the cell grid is shrunk to at most
// 2 count + 64 cells, which only makes cells larger.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>

#include "cubble_amd.h"
#include "dev_util.hpp"

namespace cubble_amd {

void set_global_error(const std::string &m);

namespace {

constexpr int kScanChunk = kBlock * 16;  // elements per scan block

struct PairScratch {
    unsigned long long *max_r_bits;  // largest radius (bit pattern of a positive double)
    int32_t *grid;                   // cell_dim[3], num_cells
    int32_t *cell_count, *cell_start, *cell_cursor, *cell_part;
    int32_t *cell_of, *members, *cnt, *off, *off_part;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

int cells_cap_of(int n) { return 2 * n + 64; }

int chunks(int m) { return (m + kScanChunk - 1) / kScanChunk; }

// Assigns the scratch arrays in one block at `base` (nullptr: size only).
size_t carve_pairs(PairScratch &P, char *base, int n) {
    size_t off = 0;
    auto take = [&](auto *&p, size_t count) {
        using T = std::remove_reference_t<decltype(*p)>;
        const size_t o = off;
        off = align256(off + std::max<size_t>(count, 1) * sizeof(T));
        p = base ? reinterpret_cast<T *>(base + o) : nullptr;
    };
    const int cc = cells_cap_of(n);
    take(P.max_r_bits, 1);
    take(P.grid, 4);
    take(P.cell_count, (size_t)cc);
    take(P.cell_start, (size_t)cc + 1);
    take(P.cell_cursor, (size_t)cc);
    take(P.cell_part, (size_t)chunks(cc) + 1);
    take(P.cell_of, (size_t)n);
    take(P.members, (size_t)n);
    take(P.cnt, (size_t)n);
    take(P.off, (size_t)n + 1);
    take(P.off_part, (size_t)chunks(n) + 1);
    return off;
}

// Coordinate wrapped into the box on a periodic axis (one period, as the
// reference's wrap).
__device__ __forceinline__ double wrap_in(const Consts &c, int k, double x) {
    if (c.wall[k]) return x;
    return x < c.lbb[k] ? x + c.interval[k] : (x > c.tfr[k] ? x - c.interval[k] : x);
}

__global__ void k_p_init(PairScratch P) {
    if (0 == threadIdx.x) *P.max_r_bits = 0ull;
}

__global__ __launch_bounds__(kBlock) void k_p_maxr(const double *r, int n, PairScratch P) {
    double m = 0.0;
    for (int i = blockIdx.x * kBlock + (int)threadIdx.x; i < n; i += gridDim.x * kBlock)
        m = fmax(m, r[i]);
    m = block_max(m);
    if (0 == threadIdx.x && m > 0.0)
        atomicMax(P.max_r_bits, (unsigned long long)__double_as_longlong(m));
}

__device__ __forceinline__ void grid_of(const Consts &c, const PairScratch &P, int cap,
                                        int cd[3]) {
    const double max_r = __longlong_as_double((long long)*P.max_r_bits);
    cell_dims(c, max_r, cap, cd);
}

__global__ __launch_bounds__(kBlock) void k_p_grid(Consts c, PairScratch P, int cells_cap) {
    int cd[3];
    grid_of(c, P, cells_cap, cd);
    const int nc = cd[0] * cd[1] * cd[2];
    if (0 == blockIdx.x && 0 == threadIdx.x) {
        P.grid[0] = cd[0];
        P.grid[1] = cd[1];
        P.grid[2] = cd[2];
        P.grid[3] = nc;
    }
    for (int q = blockIdx.x * kBlock + (int)threadIdx.x; q < nc; q += gridDim.x * kBlock)
        P.cell_count[q] = 0;
}

template <int D>
__device__ __forceinline__ int cell_of_pos(const Consts &c, const int cd[3], const double p[3]) {
    int id[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < D; ++q) {
        const double f = floor(cd[q] * (p[q] - c.lbb[q]) / c.interval[q]);
        id[q] = f < 0.0 ? 0 : (f >= (double)cd[q] ? cd[q] - 1 : (int)f);
    }
    return id[2] * cd[0] * cd[1] + id[1] * cd[0] + id[0];
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_p_cells(Consts c, const double *x, const double *y,
                                                    const double *z, int n, PairScratch P) {
    const int cd[3] = {P.grid[0], P.grid[1], P.grid[2]};
    for (int i = blockIdx.x * kBlock + (int)threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double p[3] = {wrap_in(c, 0, x[i]), wrap_in(c, 1, y[i]),
                             3 == D ? wrap_in(c, 2, z[i]) : 0.0};
        const int cid = cell_of_pos<D>(c, cd, p);
        P.cell_of[i] = cid;
        atomicAdd(P.cell_count + cid, 1);
    }
}

// Exclusive scan of v[0, m) into out[0, m] (out[m] = total), in chunks of
// kScanChunk: pass 1 sums each chunk, pass 2 scans a chunk after the sums of
// the chunks before it.  *m_ptr (device) holds m when m_ptr is set.
__global__ __launch_bounds__(kBlock) void k_p_scan_part(const int32_t *v, int m,
                                                        const int32_t *m_ptr, int32_t *part) {
    if (m_ptr) m = *m_ptr;
    const int base = blockIdx.x * kScanChunk;
    if (base >= m) return;  // uniform per block
    int s = 0;
    for (int q = base + (int)threadIdx.x; q < min(base + kScanChunk, m); q += kBlock) s += v[q];
    s = block_sum_i(s);
    if (0 == threadIdx.x) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_p_scan_out(const int32_t *v, int m,
                                                       const int32_t *m_ptr, const int32_t *part,
                                                       int32_t *out, int32_t *copy) {
    if (m_ptr) m = *m_ptr;
    const int base = blockIdx.x * kScanChunk;
    if (base >= m) return;  // uniform per block
    int acc = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += kBlock) acc += part[b];
    int off = block_sum_i(acc);
    for (int k = 0; k < kScanChunk / kBlock; ++k) {
        const int q = base + k * kBlock + (int)threadIdx.x;
        const int val = q < m ? v[q] : 0;
        int tot = 0;
        const int pre = block_excl_scan(val, tot);
        if (q < m) {
            out[q] = off + pre;
            if (copy) copy[q] = off + pre;
        }
        off += tot;
    }
    if (m - 1 < base + kScanChunk && 0 == threadIdx.x) out[m] = off;
}

__global__ __launch_bounds__(kBlock) void k_p_bin(int n, PairScratch P) {
    for (int i = blockIdx.x * kBlock + (int)threadIdx.x; i < n; i += gridDim.x * kBlock)
        P.members[atomicAdd(P.cell_cursor + P.cell_of[i], 1)] = i;
}

// Sorted, de-duplicated cells of coordinate v along one axis (periodic axes
// wrap; wall axes drop cells outside the grid).
__device__ __forceinline__ int axis_neighbours(int v, int n, bool wall, int out[3]) {
    int k = 0;
    for (int o = -1; o <= 1; ++o) {
        int w = v + o;
        if (wall) {
            if (w < 0 || w >= n) continue;
        } else {
            w = (w + n) % n;
        }
        bool dup = false;
        for (int t = 0; t < k; ++t) dup = dup || out[t] == w;
        if (!dup) out[k++] = w;
    }
    return k;
}

// Candidates j > i of bubble i over its stencil: FILL = false counts them
// into cnt[i]; FILL = true writes them at off[i] (within pair_cap) and sorts
// the bubble's group by j.
template <int D, bool FILL>
__global__ __launch_bounds__(kBlock) void k_p_walk(Consts c, const double *x, const double *y,
                                                   const double *z, const double *r, int n,
                                                   PairScratch P, int32_t *pair_i,
                                                   int32_t *pair_j, long long pair_cap) {
#pragma clang fp contract(off)
    const int cd[3] = {P.grid[0], P.grid[1], P.grid[2]};
    for (int i = blockIdx.x * kBlock + (int)threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double pi[3] = {wrap_in(c, 0, x[i]), wrap_in(c, 1, y[i]),
                              3 == D ? wrap_in(c, 2, z[i]) : 0.0};
        const double ri = r[i];
        const int cid = P.cell_of[i];
        const int cc[3] = {cid % cd[0], (cid / cd[0]) % cd[1], cid / (cd[0] * cd[1])};
        int ax[3][3], na[3] = {1, 1, 1};
        ax[2][0] = 0;
        for (int q = 0; q < D; ++q) na[q] = axis_neighbours(cc[q], cd[q], c.wall[q] != 0, ax[q]);
        const long long o0 = FILL ? (long long)P.off[i] : 0;
        int k = 0;
        for (int a = 0; a < na[2]; ++a)
            for (int b = 0; b < na[1]; ++b)
                for (int e = 0; e < na[0]; ++e) {
                    const int cell = ax[2][a] * cd[0] * cd[1] + ax[1][b] * cd[0] + ax[0][e];
                    for (int s = P.cell_start[cell]; s < P.cell_start[cell + 1]; ++s) {
                        const int j = P.members[s];
                        if (j <= i) continue;
                        const double pj[3] = {wrap_in(c, 0, x[j]), wrap_in(c, 1, y[j]),
                                              3 == D ? wrap_in(c, 2, z[j]) : 0.0};
                        double d[3];
                        wrapped_diff<D>(c, pi[0], pi[1], pi[2], pj[0], pj[1], pj[2], d);
                        const double maxd = ri + r[j] + c.skin;
                        if (!(sq_len(d) < maxd * maxd)) continue;
                        if (FILL && o0 + k < pair_cap) {
                            pair_i[o0 + k] = i;
                            pair_j[o0 + k] = j;
                        }
                        ++k;
                    }
                }
        if (!FILL) {
            P.cnt[i] = k;
            continue;
        }
        // the group's j ascending (insertion sort; groups hold a few pairs)
        const long long hi = min(o0 + k, pair_cap);
        for (long long q = o0 + 1; q < hi; ++q) {
            const int v = pair_j[q];
            long long t = q - 1;
            while (t >= o0 && pair_j[t] > v) {
                pair_j[t + 1] = pair_j[t];
                --t;
            }
            pair_j[t + 1] = v;
        }
    }
}

__global__ void k_p_total(int n, PairScratch P, int64_t *num_pairs) {
    if (0 == threadIdx.x) *num_pairs = n > 0 ? (int64_t)P.off[n] : 0;
}

int grid_blocks(int items) { return std::max(1, std::min((items + kBlock - 1) / kBlock, 8192)); }

// The whole search on `stream` with scratch P (carved for >= n bubbles).
int enqueue_pairs(const Consts &c, const double *x, const double *y, const double *z,
                  const double *r, int n, const PairScratch &P, int32_t *pair_i, int32_t *pair_j,
                  long long pair_cap, int64_t *num_pairs, hipStream_t st) {
    if (0 == n) {  // no bubbles, no pairs (and no empty launches)
        hipLaunchKernelGGL(k_p_total, dim3(1), dim3(64), 0, st, 0, P, num_pairs);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_global_error(std::string("find_pairs launch: ") + hipGetErrorString(e));
            return CUBBLE_E_HIP;
        }
        return CUBBLE_OK;
    }
    const int cc = cells_cap_of(n);
    const int gb = grid_blocks(n), gc = grid_blocks(cc);
    hipLaunchKernelGGL(k_p_init, dim3(1), dim3(64), 0, st, P);
    hipLaunchKernelGGL(k_p_maxr, dim3(gb), dim3(kBlock), 0, st, r, n, P);
    hipLaunchKernelGGL(k_p_grid, dim3(gc), dim3(kBlock), 0, st, c, P, cc);
    if (3 == c.dim) hipLaunchKernelGGL(k_p_cells<3>, dim3(gb), dim3(kBlock), 0, st, c, x, y, z, n, P);
    else hipLaunchKernelGGL(k_p_cells<2>, dim3(gb), dim3(kBlock), 0, st, c, x, y, z, n, P);
    // cell starts (and cursors) over the grid's cells; the cell count is on
    // the device, so the scans read it there and launch for the capacity
    const int32_t *ncell = P.grid + 3;
    hipLaunchKernelGGL(k_p_scan_part, dim3(chunks(cc)), dim3(kBlock), 0, st, P.cell_count, cc,
                       ncell, P.cell_part);
    hipLaunchKernelGGL(k_p_scan_out, dim3(chunks(cc)), dim3(kBlock), 0, st, P.cell_count, cc,
                       ncell, P.cell_part, P.cell_start, P.cell_cursor);
    hipLaunchKernelGGL(k_p_bin, dim3(gb), dim3(kBlock), 0, st, n, P);
    if (3 == c.dim)
        hipLaunchKernelGGL((k_p_walk<3, false>), dim3(gb), dim3(kBlock), 0, st, c, x, y, z, r, n,
                           P, pair_i, pair_j, pair_cap);
    else
        hipLaunchKernelGGL((k_p_walk<2, false>), dim3(gb), dim3(kBlock), 0, st, c, x, y, z, r, n,
                           P, pair_i, pair_j, pair_cap);
    hipLaunchKernelGGL(k_p_scan_part, dim3(chunks(n)), dim3(kBlock), 0, st, P.cnt, n,
                       (const int32_t *)nullptr, P.off_part);
    hipLaunchKernelGGL(k_p_scan_out, dim3(chunks(n)), dim3(kBlock), 0, st, P.cnt, n,
                       (const int32_t *)nullptr, P.off_part, P.off, (int32_t *)nullptr);
    if (3 == c.dim)
        hipLaunchKernelGGL((k_p_walk<3, true>), dim3(gb), dim3(kBlock), 0, st, c, x, y, z, r, n,
                           P, pair_i, pair_j, pair_cap);
    else
        hipLaunchKernelGGL((k_p_walk<2, true>), dim3(gb), dim3(kBlock), 0, st, c, x, y, z, r, n,
                           P, pair_i, pair_j, pair_cap);
    hipLaunchKernelGGL(k_p_total, dim3(1), dim3(64), 0, st, n, P, num_pairs);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_global_error(std::string("find_pairs launch: ") + hipGetErrorString(e));
        return CUBBLE_E_HIP;
    }
    return CUBBLE_OK;
}

// Arguments checked on the host; the box becomes the kernels' constants.
int check_args(const double *x, const double *y, const double *z, const double *r, int32_t n,
               const cubble_box *box, double skin, const int32_t *pair_i, const int32_t *pair_j,
               int64_t pair_cap, const int64_t *num_pairs, Consts &c) {
    auto bad = [](const char *m) {
        set_global_error(std::string("find_pairs: ") + m);
        return CUBBLE_E_INVALID;
    };
    if (!box || (box->dim != 2 && box->dim != 3)) return bad("box.dim must be 2 or 3");
    if (n < 0 || pair_cap < 0 || !(skin >= 0.0)) return bad("negative count, capacity or skin");
    if (!num_pairs) return bad("num_pairs is null");
    if (n > 0 && (!x || !y || !r || (3 == box->dim && !z)))
        return bad("a position or radius array is null");
    if (pair_cap > 0 && (!pair_i || !pair_j)) return bad("pair arrays are null");
    std::memset(&c, 0, sizeof(c));
    c.dim = box->dim;
    for (int k = 0; k < 3; ++k) {
        const bool used = k < box->dim;
        c.wall[k] = used ? (box->wall[k] != 0) : 0;
        c.lbb[k] = used ? box->lbb[k] : 0.0;
        c.tfr[k] = used ? box->tfr[k] : 0.0;
        c.interval[k] = c.tfr[k] - c.lbb[k];
        if (used && !(c.interval[k] > 0.0)) return bad("box.tfr must exceed box.lbb");
    }
    c.skin = skin;
    return CUBBLE_OK;
}

}  // namespace

// Arguments checked, scratch taken (from `scratch`, `scratch_bytes` long, or
// when that is null with hipMallocAsync / hipFreeAsync on the stream), the
// search enqueued.
int find_pairs_run(const double *x, const double *y, const double *z, const double *r,
                   int32_t count, const cubble_box *box, double skin, int32_t *pair_i,
                   int32_t *pair_j, int64_t pair_cap, int64_t *num_pairs, void *scratch,
                   size_t scratch_bytes, hipStream_t st) {
    Consts c;
    int rc = check_args(x, y, z, r, count, box, skin, pair_i, pair_j, pair_cap, num_pairs, c);
    if (rc) return rc;
    PairScratch P;
    const size_t bytes = carve_pairs(P, nullptr, count);
    if (scratch) {
        if (bytes > scratch_bytes) {
            set_global_error("find_pairs: " + std::to_string(count) +
                             " bubbles need " + std::to_string(bytes) +
                             " bytes of scratch, more than the " +
                             std::to_string(scratch_bytes) + " available");
            return CUBBLE_E_CAPACITY;
        }
        carve_pairs(P, static_cast<char *>(scratch), count);
        return enqueue_pairs(c, x, y, z, r, count, P, pair_i, pair_j, pair_cap, num_pairs, st);
    }
    void *mem = nullptr;
    hipError_t e = hipMallocAsync(&mem, bytes, st);
    if (e != hipSuccess) {
        set_global_error(std::string("find_pairs: hipMallocAsync: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? CUBBLE_E_NOMEM : CUBBLE_E_HIP;
    }
    carve_pairs(P, static_cast<char *>(mem), count);
    rc = enqueue_pairs(c, x, y, z, r, count, P, pair_i, pair_j, pair_cap, num_pairs, st);
    e = hipFreeAsync(mem, st);
    if (!rc && e != hipSuccess) {
        set_global_error(std::string("find_pairs: hipFreeAsync: ") + hipGetErrorString(e));
        rc = CUBBLE_E_HIP;
    }
    return rc;
}

size_t find_pairs_scratch_bytes(int32_t count) {
    PairScratch P;
    return carve_pairs(P, nullptr, count < 0 ? 0 : count);
}

}  // namespace cubble_amd

using namespace cubble_amd;

extern "C" {

int cubble_find_pairs(const double *x, const double *y, const double *z, const double *r,
                      int32_t count, const cubble_box *box, double skin, int32_t *pair_i,
                      int32_t *pair_j, int64_t pair_cap, int64_t *num_pairs, void *stream) {
    return find_pairs_run(x, y, z, r, count, box, skin, pair_i, pair_j, pair_cap, num_pairs,
                          nullptr, 0, static_cast<hipStream_t>(stream));
}

}  // extern "C"
