// remove_kernels.hip -- removeBubbles (reference src/simulator.cu:180-190).
//
//   k_rm_emit  : blocks holding marked bubbles write them, ascending, into the
//                deletion list at their prefix offset (block counts come from
//                the corrector's partials)
//   k_rm_plan  : one block per foam plans the moves -- the k-th hole below the
//                new count receives the k-th surviving bubble counted from the
//                back (reference swapDataCountPairs, src/kernels.cu:775-857) --
//                and the volume-fix multiplier M / S_V^(1/D)
//                (src/kernels.cu:861-867)
//   k_rm_mark  : marks the lists that change: those of the neighbours of every
//                deleted bubble and of every moving tail survivor
//   k_rm_apply : every slot below the new count applies the volume fix; a hole
//                pulls its tail survivor's state and list, a marked bubble
//                filters its own list -- deleted neighbours dropped, moved
//                ones mapped to their new slots, kept ascending (reference
//                addVolumeFixPairs, src/kernels.cu:859-936)
//   k_rm_fin   : one block per foam commits the new count
//
// The reference appends the deletion list with an atomic counter, so its
// hole / survivor pairing depends on thread timing; a sorted list makes the
// pairing deterministic.  Its race-prone pair swap-from-the-back
// (src/kernels.cu:873-909) is replaced by per-bubble list filtering, done
// only for the few lists a removal touches.
#include "dev_util.hpp"
#include "launch.hpp"

namespace cubble_amd {

namespace {

constexpr int kFilter = 8;  // list entries loaded together when walking a list

__device__ __forceinline__ void emit_block(Foam *F, int lb) {
    const Scratch S = F->s;
    const int n = F->d.count, ipt = F->ipt;
    const int cnt = S.part_i[lb * kPartI + 0];
    if (cnt == 0) return;  // uniform per block
    int acc = 0;
    for (int b = threadIdx.x; b < lb; b += kBlock) acc += S.part_i[b * kPartI + 0];
    int off = block_sum_i(acc);
    const int base = lb * kBlock * ipt;
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        const int f = (i < n && S.del[i]) ? 1 : 0;
        int tot = 0;
        const int pre = block_excl_scan(f, tot);
        if (f) S.del_list[off + pre] = i;
        off += tot;
    }
}

// Items of the mark kernel: the deleted bubbles, then the moving survivors
// (at most the old count, so the bubble-space blocks cover them).
__device__ __forceinline__ void mark_block(Foam *F, int lb) {
    const Scratch S = F->s;
    const int nd = F->d.n_del, items = nd + F->d.n_holes;
    const int nnew = F->d.count - nd;
    const size_t cap = (size_t)F->cap;
    const int base = lb * kBlock * F->ipt;
    if (base >= items) return;  // uniform per block
    for (int k = 0; k < F->ipt; ++k) {
        const int t = base + k * kBlock + (int)threadIdx.x;
        if (t >= items) break;
        const int s = t < nd ? S.del_list[t] : S.plan_src[t - nd];
        const int m = S.nn[s];
        // list entries loaded kFilter at a time (independent loads)
        for (int q0 = 0; q0 < m; q0 += kFilter) {
            int js[kFilter];
#pragma unroll
            for (int u = 0; u < kFilter; ++u)
                js[u] = q0 + u < m ? S.nbr[(size_t)(q0 + u) * cap + s] : nnew;
#pragma unroll
            for (int u = 0; u < kFilter; ++u)
                if (js[u] < nnew) S.dirty[js[u]] = 1;  // same value from every writer
        }
    }
}

template <int D, typename V>
__device__ __forceinline__ int apply_range(const Roles &R, const Scratch &S, size_t cap, int base,
                                           int ipt, int nnew, int h, double vm) {
    int sum_nn = 0;
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= nnew) break;
        const bool hole = S.del[i] != 0;
        const bool marked = S.dirty[i] != 0;
        int src = i;
        if (hole) {
            int lo = 0, hi = h;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (S.del_list[mid] < i) lo = mid + 1;
                else hi = mid;
            }
            src = S.plan_src[lo];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (2 == q && 2 == D) continue;
                R.x[q][i] = R.x[q][src];
                field<V>(R.v[q])[i] = field<V>(R.v[q])[src];
                field<V>(R.vo[q])[i] = field<V>(R.vo[q])[src];
                field<V>(R.saved[q])[i] = field<V>(R.saved[q])[src];
            }
            R.path[i] = R.path[src];
            field<V>(R.err)[i] = field<V>(R.err)[src];
#pragma unroll
            for (int q = 0; q < D; ++q) R.wc[q][i] = R.wc[q][src];
            R.idx[i] = R.idx[src];
            S.deg[i] = S.deg[src];  // the reference moves numNeighbors unchanged
        }
        int out = S.nn[src];
        if (hole || marked) {
            const int m = out;
            out = 0;
            int last = -1;  // largest entry written so far
            // entries and their deletion marks are loaded kFilter at a time
            // (independent loads); a chunk is read before any of its slots is
            // written, and writes stay at positions <= the entries consumed,
            // so the filter also runs in place (src == i)
            for (int q0 = 0; q0 < m; q0 += kFilter) {
                int js[kFilter];
                uint8_t dl[kFilter];
#pragma unroll
                for (int u = 0; u < kFilter; ++u)
                    js[u] = q0 + u < m ? S.nbr[(size_t)(q0 + u) * cap + src] : -1;
#pragma unroll
                for (int u = 0; u < kFilter; ++u) dl[u] = js[u] >= 0 ? S.del[js[u]] : 1;
#pragma unroll
                for (int u = 0; u < kFilter; ++u) {
                    if (dl[u]) continue;
                    const int j = js[u];
                    const int jj = j >= nnew ? S.moved_to[j - nnew] : j;
                    if (jj > last) {  // still ascending: append
                        S.nbr[(size_t)out * cap + i] = jj;
                        last = jj;
                    } else {  // a moved bubble's new index: insertion keeps the order
                        int pos = out;
                        while (pos > 0) {
                            const int prev = S.nbr[(size_t)(pos - 1) * cap + i];
                            if (prev <= jj) break;
                            S.nbr[(size_t)pos * cap + i] = prev;
                            --pos;
                        }
                        S.nbr[(size_t)pos * cap + i] = jj;
                    }
                    ++out;
                }
            }
            S.nn[i] = out;
            if (marked) S.dirty[i] = 0;
        }
        R.x[3][i] = R.x[3][i] * vm;
        sum_nn += out;
    }
    return sum_nn;
}

__device__ __forceinline__ void apply_block(Foam *F, int lb) {
    const Roles R = F->r;
    const Scratch S = F->s;
    const int n = F->d.count, ipt = F->ipt;
    const int nnew = n - F->d.n_del, h = F->d.n_holes;
    const double vm = F->d.vol_fix;
    const size_t cap = (size_t)F->cap;
    const int base = lb * kBlock * ipt;
    int sum_nn;
    if (CUBBLE_PREC_MIXED == F->precision)  // FP32 velocities
        sum_nn = 3 == F->c.dim ? apply_range<3, float>(R, S, cap, base, ipt, nnew, h, vm)
                               : apply_range<2, float>(R, S, cap, base, ipt, nnew, h, vm);
    else
        sum_nn = 3 == F->c.dim ? apply_range<3, double>(R, S, cap, base, ipt, nnew, h, vm)
                               : apply_range<2, double>(R, S, cap, base, ipt, nnew, h, vm);
    sum_nn = block_sum_i(sum_nn);
    if (threadIdx.x == 0) S.part_i[lb * kPartI + 1] = sum_nn;
}

}  // namespace

__global__ __launch_bounds__(kBlock) void k_rm_emit(Foam *foams, const int2 *bt) {
    const Ctx cx = bubble_ctx(foams, bt);
    if (cx.F->d.need_remove) emit_block(cx.F, cx.lb);
}

__global__ __launch_bounds__(kBlock) void k_rm_plan(Foam *foams) {
    Foam *F = foams + blockIdx.x;
    if (!F->d.need_remove) return;
    const Scratch S = F->s;
    const int n = F->d.count, nd = F->d.n_del, nnew = n - nd;
    int below = 0;
    for (int k = threadIdx.x; k < nd; k += kBlock) below += S.del_list[k] < nnew ? 1 : 0;
    const int h = block_sum_i(below);
    int got = 0;
    for (int top = n - 1; got < h && top >= nnew; top -= kBlock) {
        const int i = top - (int)threadIdx.x;
        const int f = (i >= nnew && !S.del[i]) ? 1 : 0;
        int tot = 0;
        const int pre = block_excl_scan(f, tot);
        if (f && got + pre < h) {
            S.plan_src[got + pre] = i;
            S.moved_to[i - nnew] = S.del_list[got + pre];
        }
        got += tot;
    }
    if (threadIdx.x == 0) {
        const double sv = F->d.s_v;
        double vm = 3 == F->c.dim ? 1.0 / det_cbrt(sv) : 1.0 / sqrt(sv);
        vm *= F->c.vol_mul;
        F->d.vol_fix = vm;
        F->d.n_holes = h;
    }
}

__global__ __launch_bounds__(kBlock) void k_rm_mark(Foam *foams, const int2 *bt) {
    const Ctx cx = cell_ctx(foams, bt);  // items in plain block order
    if (cx.F->d.need_remove) mark_block(cx.F, cx.lb);
}

__global__ __launch_bounds__(kBlock) void k_rm_apply(Foam *foams, const int2 *bt) {
    const Ctx cx = bubble_ctx(foams, bt);
    if (cx.F->d.need_remove) apply_block(cx.F, cx.lb);
}

__global__ __launch_bounds__(kBlock) void k_rm_fin(Foam *foams) {
    Foam *F = foams + blockIdx.x;
    if (!F->d.need_remove) return;
    const int nb = F->nblk;
    int s = 0;
    for (int b = threadIdx.x; b < nb; b += kBlock) s += F->s.part_i[b * kPartI + 1];
    s = block_sum_i(s);
    if (threadIdx.x == 0) {
        F->d.count = F->d.count - F->d.n_del;
        F->d.removals += 1;
        F->d.num_pairs = s / 2;
        F->d.need_remove = 0;
        if (F->d.count_stop > 0 && F->d.count <= F->d.count_stop) F->d.halt = 1;
    }
}

void launch_remove(const Grid &g) {
    hipLaunchKernelGGL(k_rm_emit, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams, g.bt_bub);
    hipLaunchKernelGGL(k_rm_plan, dim3(g.nfoams), dim3(kBlock), 0, g.stream, g.foams);
    hipLaunchKernelGGL(k_rm_mark, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams, g.bt_bub);
    hipLaunchKernelGGL(k_rm_apply, dim3(g.nb_bub), dim3(kBlock), 0, g.stream, g.foams, g.bt_bub);
    hipLaunchKernelGGL(k_rm_fin, dim3(g.nfoams), dim3(kBlock), 0, g.stream, g.foams);
}

}  // namespace cubble_amd
