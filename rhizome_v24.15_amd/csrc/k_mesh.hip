// k_mesh.hip -- mesh intake: triangle build with the degenerate filter and
// the vertex bounding box.
//
// Triangle build follows dev_try_build_cuda_triangle (core/math_utils.cuh:
// 484-514): out-of-range indices and |(v1-v0)x(v2-v0)|^2 <= 1e-12 are dropped,
// the normal is the unit cross product.  The reference compacts with an
// atomicAdd slot (cuda_rasterizer.cu:1334), which scrambles triangle order
// from run to run; here a flag + exclusive scan keeps index order, so the
// FP32 winding sums downstream are reproducible.
#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

__device__ __forceinline__ bool build_one(const float* __restrict__ v, int64_t nv, const uint32_t* __restrict__ idx,
                                          int64_t t, float out[12]) {
    uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
    if ((int64_t)i0 >= nv || (int64_t)i1 >= nv || (int64_t)i2 >= nv) return false;
    float ax = v[3 * (int64_t)i0], ay = v[3 * (int64_t)i0 + 1], az = v[3 * (int64_t)i0 + 2];
    float bx = v[3 * (int64_t)i1], by = v[3 * (int64_t)i1 + 1], bz = v[3 * (int64_t)i1 + 2];
    float cx = v[3 * (int64_t)i2], cy = v[3 * (int64_t)i2 + 1], cz = v[3 * (int64_t)i2 + 2];
    float e1x = bx - ax, e1y = by - ay, e1z = bz - az;
    float e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
    float nx = e1y * e2z - e2y * e1z;
    float ny = e1z * e2x - e2z * e1x;
    float nz = e1x * e2y - e2x * e1y;
    float len2 = dot3(nx, ny, nz, nx, ny, nz);
    if (!(len2 > 1e-12f)) return false;
    float r = 1.0f / sqrtf(len2);
    out[0] = ax;
    out[1] = ay;
    out[2] = az;
    out[3] = bx;
    out[4] = by;
    out[5] = bz;
    out[6] = cx;
    out[7] = cy;
    out[8] = cz;
    out[9] = nx * r;
    out[10] = ny * r;
    out[11] = nz * r;
    return true;
}

__global__ void tri_valid_kernel(const float* __restrict__ v, int64_t nv, const uint32_t* __restrict__ idx,
                                 int64_t nt, uint32_t* __restrict__ valid) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    float tmp[12];
    valid[t] = build_one(v, nv, idx, t, tmp) ? 1u : 0u;
}

__global__ void tri_emit_kernel(const float* __restrict__ v, const uint32_t* __restrict__ idx, int64_t nt,
                                const uint32_t* __restrict__ valid, const int64_t* __restrict__ pos,
                                float* __restrict__ out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt || !valid[t]) return;
    float tmp[12];
    build_one(v, INT64_MAX, idx, t, tmp);
    float* o = out + 12 * pos[t];
    for (int j = 0; j < 12; ++j) o[j] = tmp[j];
}

// Per-block min/max, then one block folds the partials (deterministic).
__global__ void bbox_partial_kernel(const float* __restrict__ v, int64_t nv, float* __restrict__ partial) {
    __shared__ float sh[6][256];
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x)
        for (int j = 0; j < 3; ++j) {
            float x = v[3 * i + j];
            if (x < lo[j]) lo[j] = x;
            if (x > hi[j]) hi[j] = x;
        }
    for (int j = 0; j < 3; ++j) {
        sh[j][threadIdx.x] = lo[j];
        sh[3 + j][threadIdx.x] = hi[j];
    }
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int j = 0; j < 3; ++j) {
                sh[j][threadIdx.x] = fminf(sh[j][threadIdx.x], sh[j][threadIdx.x + st]);
                sh[3 + j][threadIdx.x] = fmaxf(sh[3 + j][threadIdx.x], sh[3 + j][threadIdx.x + st]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < 6; ++j) partial[6 * blockIdx.x + j] = sh[j][0];
}

__global__ void bbox_final_kernel(const float* __restrict__ partial, int nb, float* __restrict__ out6) {
    if (threadIdx.x != 0) return;
    float r[6] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f};
    for (int b = 0; b < nb; ++b)
        for (int j = 0; j < 3; ++j) {
            r[j] = fminf(r[j], partial[6 * b + j]);
            r[3 + j] = fmaxf(r[3 + j], partial[6 * b + 3 + j]);
        }
    for (int j = 0; j < 6; ++j) out6[j] = r[j];
}

}  // namespace

hipError_t launch_tri_valid(const float* verts, int64_t nv, const uint32_t* idx, int64_t nt, uint32_t* valid,
                            hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    tri_valid_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(verts, nv, idx, nt, valid);
    return hipGetLastError();
}

hipError_t launch_tri_emit(const float* verts, const uint32_t* idx, int64_t nt, const uint32_t* valid,
                           const int64_t* pos, float* out, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    tri_emit_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(verts, idx, nt, valid, pos, out);
    return hipGetLastError();
}

int bbox_blocks(int64_t nv) {
    int64_t b = (nv + 255) / 256;
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t launch_bbox(const float* verts, int64_t nv, float* partial, int nblocks, float* out6, hipStream_t s) {
    bbox_partial_kernel<<<nblocks, 256, 0, s>>>(verts, nv, partial);
    bbox_final_kernel<<<1, 64, 0, s>>>(partial, nblocks, out6);
    return hipGetLastError();
}

}  // namespace rhz
