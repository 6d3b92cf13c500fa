// pairs_noalloc.hip -- cubble_find_pairs_noalloc: the pair search of
// cubble_find_pairs (pairs.hip, same kernels) with its scratch in a fixed
// device array instead of stream-ordered allocations, for callers that cannot
// allocate device memory: code under HIP graph capture, serving loops that
// must not call hipMalloc.
//
// The array is a file-scope __device__ variable: every device the library
// runs on holds its own copy, in the code object's data, until the process
// ends.  In host code its name is a host-side stand-in, not a device address:
// the device address of the current device's copy comes from
// hipGetSymbolAddress, taken at the first call on each device and kept per
// device.  All calls on one device share that copy, so calls that overlap on
// one device race; the kernels clear what must start at zero (the radius
// maximum, the cell counts), whatever the previous call left.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "cubble_amd.h"

namespace cubble_amd {

void set_global_error(const std::string &m);
int find_pairs_run(const double *x, const double *y, const double *z, const double *r,
                   int32_t count, const cubble_box *box, double skin, int32_t *pair_i,
                   int32_t *pair_j, int64_t pair_cap, int64_t *num_pairs, void *scratch,
                   size_t scratch_bytes, hipStream_t st);

namespace {

constexpr size_t kScratchBytes = (size_t)CUBBLE_FIND_PAIRS_NOALLOC_SCRATCH_MIB << 20;

__device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

}  // namespace
}  // namespace cubble_amd

using namespace cubble_amd;

extern "C" {

int cubble_find_pairs_noalloc(const double *x, const double *y, const double *z, const double *r,
                              int32_t count, const cubble_box *box, double skin, int32_t *pair_i,
                              int32_t *pair_j, int64_t pair_cap, int64_t *num_pairs,
                              void *stream) {
    if (count > CUBBLE_FIND_PAIRS_NOALLOC_MAX) {
        set_global_error("find_pairs_noalloc: " + std::to_string(count) +
                         " bubbles exceed the limit of " +
                         std::to_string(CUBBLE_FIND_PAIRS_NOALLOC_MAX));
        return CUBBLE_E_CAPACITY;
    }
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_global_error("find_pairs_noalloc: no usable current device");
        return CUBBLE_E_HIP;
    }
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        if (!g_scratch_ptr[dev]) {
            void *p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                set_global_error(std::string("find_pairs_noalloc: hipGetSymbolAddress: ") +
                                 hipGetErrorString(e));
                return CUBBLE_E_HIP;
            }
            g_scratch_ptr[dev] = p;
        }
        base = g_scratch_ptr[dev];
    }
    return find_pairs_run(x, y, z, r, count, box, skin, pair_i, pair_j, pair_cap, num_pairs, base,
                          kScratchBytes, static_cast<hipStream_t>(stream));
}

}  // extern "C"
