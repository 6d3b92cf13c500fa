// observe_kernels.hip -- the observables the reference's integration loop
// reports at each printed row and snapshot, reduced on the device:
//
//   totalEnergy   (src/simulator.cu:39-58, kernel potentialEnergy
//                  src/kernels.cu:941-956): per bubble, the sum over its
//                  overlapping neighbours of (r_i + r_j - |d_ij|)^2, with
//                  d_ij the minimum-image difference; the total is the sum
//                  over bubbles (each pair counted from both ends)
//   totalVolume   (src/simulator.cu:362-379, kernel calculateVolumes
//                  src/kernels.cu:1013-1024)
//   getAvg        (src/simulator.cu:1036-1050): means of r, path, dxdt,
//                  dydt, dzdt, drdt
//
// k_observe walks the full neighbour lists (both directions), so every
// bubble's energy is a sum over its own list and needs no atomics; block
// partials go to the foam's partial records and k_observe_fin reduces them in
// block order, so the totals are deterministic.
#include <algorithm>

#include "cubble_amd.h"
#include "dev_util.hpp"
#include "launch.hpp"

namespace cubble_amd {

namespace {

enum Obs { kObsEnergy, kObsVolume, kObsR, kObsPath, kObsV0 };  // kObsV0 + q: mean v[q]

template <int D, typename V>
__global__ __launch_bounds__(kBlock) void k_observe(Foam *F, double *energy) {
    const Consts c = F->c;
    const Roles R = F->r;
    const Scratch S = F->s;
    const int n = F->d.count, ipt = F->ipt;
    const size_t cap = (size_t)F->cap;
    const int base = blockIdx.x * kBlock * ipt;
    double acc[kPartD];
#pragma unroll
    for (int q = 0; q < kPartD; ++q) acc[q] = 0.0;
    for (int k = 0; k < ipt; ++k) {
        const int i = base + k * kBlock + (int)threadIdx.x;
        if (i >= n) break;
        const double xi = R.x[0][i], yi = R.x[1][i], zi = 3 == D ? R.x[2][i] : 0.0;
        const double ri = R.x[3][i];
        double e = 0.0;
        const int m = S.nn[i];
        for (int q = 0; q < m; ++q) {
            const int j = S.nbr[(size_t)q * cap + i];
            double d[3];
            wrapped_diff<D>(c, xi, yi, zi, R.x[0][j], R.x[1][j], 3 == D ? R.x[2][j] : 0.0, d);
            double o = ri + R.x[3][j] - sqrt(sq_len(d));
            if (o > 0) e += o * o;
        }
        if (energy) energy[i] = e;
        double vol = ri * ri * kPi;
        if (3 == D) vol *= ri * 1.33333333333333333333333333;
        acc[kObsEnergy] += e;
        acc[kObsVolume] += vol;
        acc[kObsR] += ri;
        acc[kObsPath] += R.path[i];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (2 != q || 3 == D) acc[kObsV0 + q] += (double)field<V>(R.v[q])[i];
    }
#pragma unroll
    for (int q = 0; q < kPartD; ++q) {
        const double s = block_sum(acc[q]);
        if (threadIdx.x == 0) S.part[blockIdx.x * kPartD + q] = s;
    }
}

__global__ __launch_bounds__(kBlock) void k_observe_fin(Foam *F, int nb, double *out) {
    const double *part = F->s.part;
#pragma unroll
    for (int q = 0; q < kPartD; ++q) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += kBlock) s += part[b * kPartD + q];
        s = block_sum(s);
        if (threadIdx.x == 0) out[q] = s;
    }
}

// Box rescale (reference transformPositions, src/kernels.cu:970-990): each
// coordinate to its relative position in the old box, then into the new one.
__global__ __launch_bounds__(kBlock) void k_rescale(Foam *F, Consts old_c, Consts new_c) {
    const Roles R = F->r;
    const int n = F->d.count, D = old_c.dim;
    for (int i = blockIdx.x * kBlock + (int)threadIdx.x; i < n; i += gridDim.x * kBlock)
        for (int a = 0; a < D; ++a) {
            const double rel = (R.x[a][i] - old_c.lbb[a]) / old_c.interval[a];
            R.x[a][i] = new_c.interval[a] * rel + new_c.lbb[a];
        }
}

}  // namespace

void launch_rescale(const Foam &h, Foam *dev, const Consts &new_c, hipStream_t stream) {
    const int nb = std::max(1, std::min(4096, (h.d.count + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_rescale, dim3(nb), dim3(kBlock), 0, stream, dev, h.c, new_c);
}

void launch_observe(const Foam &h, Foam *dev, double *out, double *energy, hipStream_t stream) {
    const int nb = std::max(1, std::min(h.nblk, (h.d.count + kBlock * h.ipt - 1) / (kBlock * h.ipt)));
    const bool mixed = CUBBLE_PREC_MIXED == h.precision;
    if (3 == h.c.dim) {
        if (mixed) hipLaunchKernelGGL((k_observe<3, float>), dim3(nb), dim3(kBlock), 0, stream, dev, energy);
        else hipLaunchKernelGGL((k_observe<3, double>), dim3(nb), dim3(kBlock), 0, stream, dev, energy);
    } else {
        if (mixed) hipLaunchKernelGGL((k_observe<2, float>), dim3(nb), dim3(kBlock), 0, stream, dev, energy);
        else hipLaunchKernelGGL((k_observe<2, double>), dim3(nb), dim3(kBlock), 0, stream, dev, energy);
    }
    hipLaunchKernelGGL(k_observe_fin, dim3(1), dim3(kBlock), 0, stream, dev, nb, out);
}

}  // namespace cubble_amd
