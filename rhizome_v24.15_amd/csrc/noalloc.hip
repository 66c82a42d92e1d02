// noalloc.hip -- rhz_stage_winding_voxels_noalloc: exact winding numbers at
// lattice voxels with no device allocation.
//
// The kernels are those of rhz_stage_winding_voxels (split partial sums,
// compensated FP32, binary64 recompute near |w| = 1/2); their scratch -- the
// recompute counter, the recompute list and the split partials -- lives in
// the file-scope device array `scratch` below instead of the library's cached
// workspace.  Every GPU the library runs on holds its own copy of the array
// (24 MiB + 256 B) until the process ends.  Its device address is taken once
// per device with hipGetSymbolAddress; host code uses that address, never
// the array's name (which on the host is only a stand-in for the symbol).
// The recompute count is read on the device, so the call enqueues its work
// and returns without waiting.  Calls that overlap on one GPU share the array
// and race; callers serialise them (one stream, or their own ordering).
#include <hip/hip_runtime.h>

#include <mutex>

#include "rhizome_sdf.h"
#include "rhz_internal.hpp"

namespace rhz {
namespace {

constexpr int64_t kMaxVoxels = int64_t(1) << 20;
constexpr size_t kCounterBytes = 256;
constexpr size_t kListBytes = size_t(kMaxVoxels) * sizeof(int64_t);  // 8 MiB
constexpr size_t kPartBytes = size_t(16) << 20;  // split partials of up to 2^20 voxels: < 16 MiB
constexpr size_t kScratchBytes = kCounterBytes + kListBytes + kPartBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[64] = {};

__global__ void zero_counter_kernel(unsigned long long* c) { *c = 0ull; }

// The device address of `scratch` on the current device (once per device).
int scratch_address(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return set_error_message(RHZ_ERR_NODEV, "rhz_stage_winding_voxels_noalloc: no current HIP device");
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return set_error_message(RHZ_ERR_HIP, "rhz_stage_winding_voxels_noalloc: hipGetSymbolAddress failed");
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return RHZ_OK;
}

}  // namespace
}  // namespace rhz

extern "C" int rhz_stage_winding_voxels_noalloc(const rhz_triangle* d_tris, int64_t n_tris, float h,
                                                const int32_t min_c[3], const int32_t dims[3],
                                                const int64_t* d_voxels, int64_t n_voxels, int32_t precision,
                                                double* d_w, void* stream) {
    using namespace rhz;
    if (!min_c || !dims || n_tris < 0 || (n_tris > 0 && !d_tris) || n_voxels < 0 ||
        (n_voxels > 0 && (!d_voxels || !d_w)) || (precision != RHZ_WINDING_FP32 && precision != RHZ_WINDING_FP32_KAHAN))
        return set_error_message(RHZ_ERR_ARG, "rhz_stage_winding_voxels_noalloc: invalid argument");
    if (n_voxels > kMaxVoxels || winding_scratch_bytes(n_voxels, n_tris) > kPartBytes)
        return set_error_message(RHZ_ERR_ARG,
                                 "rhz_stage_winding_voxels_noalloc: more than 2^20 voxels (the fixed scratch's limit)");
    if (n_voxels == 0) return RHZ_OK;
    Grid g;
    for (int j = 0; j < 3; ++j) {
        g.mn[j] = min_c[j];
        g.dims[j] = dims[j];
    }
    g.h = h;
    void* base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    unsigned char* b = static_cast<unsigned char*>(base);
    unsigned long long* counter = reinterpret_cast<unsigned long long*>(b);
    int64_t* list = reinterpret_cast<int64_t*>(b + kCounterBytes);
    void* part = b + kCounterBytes + kListBytes;
    rhz_sdf_params p;
    rhz_sdf_params_default(&p);
    hipStream_t s = (hipStream_t)stream;
    const float* tris = reinterpret_cast<const float*>(d_tris);
    // the counter keeps the previous call's value: clear it with a kernel
    zero_counter_kernel<<<1, 1, 0, s>>>(counter);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = launch_winding(tris, n_tris, g, d_voxels, 0, n_voxels, precision, p.fp64_band, d_w, part, list, counter, s);
    if (e == hipSuccess && precision != RHZ_WINDING_FP32)
        e = launch_winding_fp64_count(tris, n_tris, g, d_voxels, 0, list, counter, n_voxels, d_w, s);
    if (e != hipSuccess) return set_error_message(RHZ_ERR_HIP, hipGetErrorString(e));
    return RHZ_OK;
}
