// k_mc.hip -- marching cubes over a signed field (CudaRasterizer::
// extract_iso_surface, engine/mesh/cuda_marching_cubes.cu:312-556, 829-962)
// and the interior-voxel count (compute_sdf_interior_volume_gpu,
// ui/viewport/cuda_viewport_gl.cu:84-110, 702-740).
//
// Semantics kept from the reference: border voxels are raised to iso + h
// before classification (its seal pass; here applied on read, the input is not
// modified); a corner is inside when value - iso <= 0; one vertex per cut
// lattice edge, interpolated with t = clamp(v0 / (v0 - v1), 0.001, 0.999)
// (0.5 when |v0 - v1| <= 1e-7) and shared by every cube around the edge;
// triangles from the standard table, emitted (e0, e2, e1).
//
// Output order is deterministic (the reference hands out vertex and triangle
// slots with atomic counters): vertices voxel by voxel in index order, x-, y-,
// then z-edge; triangles cube by cube in table order.  Passes over the field:
// inside bytes (sealed value - iso <= 0, one byte per voxel), classify
// (cut-edge mask and triangle count per voxel, block totals),
// block prefix sums, then emission, where a block recomputes its in-block
// prefix in LDS and a vertex id is its owner voxel's offset plus its rank in
// the owner's mask.
#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

#define MC_TABLE_QUALIFIER __constant__
#include "mc_table.inc"
#undef MC_TABLE_QUALIFIER

constexpr int kMcBlock = 256;

// cube corner offsets (x, y, z) and, per cube edge, the corner that owns the
// lattice edge and its axis
__constant__ int8_t c_corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                      {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__constant__ int8_t c_owner[12][2] = {{0, 0}, {1, 1}, {3, 0}, {0, 1}, {4, 0}, {5, 1},
                                      {7, 0}, {4, 1}, {0, 2}, {1, 2}, {2, 2}, {3, 2}};

struct McField {
    const float* g;
    int dx, dy, dz;
    float iso, seal;
    int64_t n;
};

__device__ __forceinline__ float mc_val(const McField& F, int x, int y, int z) {
    float v = F.g[((int64_t)z * F.dy + y) * F.dx + x];
    if ((x == 0 || x == F.dx - 1 || y == 0 || y == F.dy - 1 || z == 0 || z == F.dz - 1) && v < F.seal) v = F.seal;
    return v;
}

__device__ __forceinline__ bool mc_inside(const McField& F, float v) { return v - F.iso <= 0.0f; }

__device__ __forceinline__ void mc_coords(const McField& F, int64_t v, int& x, int& y, int& z) {
    x = (int)(v % F.dx);
    const int64_t r = v / F.dx;
    y = (int)(r % F.dy);
    z = (int)(r / F.dy);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
    return v;
}

// Pass 0: one byte per voxel, 1 when the sealed value is inside.
__global__ __launch_bounds__(kMcBlock) void mc_inside_kernel(McField F, uint8_t* __restrict__ in) {
    for (int64_t v = (int64_t)blockIdx.x * kMcBlock + threadIdx.x; v < F.n; v += (int64_t)gridDim.x * kMcBlock) {
        int x, y, z;
        mc_coords(F, v, x, y, z);
        in[v] = mc_inside(F, mc_val(F, x, y, z)) ? 1 : 0;
    }
}

// Cube case of the cube at v from the inside bytes.
__device__ __forceinline__ int mc_case_bytes(const uint8_t* __restrict__ in, int64_t v, int64_t ys, int64_t zs) {
    return (int)in[v] | ((int)in[v + 1] << 1) | ((int)in[v + ys + 1] << 2) | ((int)in[v + ys] << 3) |
           ((int)in[v + zs] << 4) | ((int)in[v + zs + 1] << 5) | ((int)in[v + zs + ys + 1] << 6) |
           ((int)in[v + zs + ys] << 7);
}

// Exclusive prefix sum of `val` over the block (kMcBlock threads); *total
// receives the block sum.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t val, uint32_t* sh, uint32_t* total) {
    const int t = threadIdx.x;
    sh[t] = val;
    __syncthreads();
    for (int off = 1; off < kMcBlock; off <<= 1) {
        const uint32_t add = (t >= off) ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const uint32_t incl = sh[t];
    *total = sh[kMcBlock - 1];
    __syncthreads();
    return incl - val;
}

// Pass 1: info[v] = cut mask (bit a: the edge from v along axis a) | triangle
// count of the cube at v << 3; per-block totals of vertices and triangles.
__global__ __launch_bounds__(kMcBlock) void mc_classify_kernel(McField F, const uint8_t* __restrict__ in,
                                                               uint8_t* __restrict__ info,
                                                               uint32_t* __restrict__ blk_v,
                                                               uint32_t* __restrict__ blk_t) {
    __shared__ uint32_t sv, st;
    if (threadIdx.x == 0) sv = st = 0;
    __syncthreads();
    const int64_t v = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
    uint32_t nvert = 0, ntri = 0;
    if (v < F.n) {
        int x, y, z;
        mc_coords(F, v, x, y, z);
        const int64_t ys = F.dx, zs = (int64_t)F.dx * F.dy;
        const uint8_t s0 = in[v];
        int mask = 0;
        if (x + 1 < F.dx && s0 != in[v + 1]) mask |= 1;
        if (y + 1 < F.dy && s0 != in[v + ys]) mask |= 2;
        if (z + 1 < F.dz && s0 != in[v + zs]) mask |= 4;
        if (x + 1 < F.dx && y + 1 < F.dy && z + 1 < F.dz) {
            const int c = mc_case_bytes(in, v, ys, zs);
            ntri = (uint32_t)(kMcTriStart[c + 1] - kMcTriStart[c]) / 3u;
        }
        nvert = (uint32_t)__popc(mask);
        info[v] = (uint8_t)(mask | (int)(ntri << 3));
    }
    const uint32_t wv = wave_sum(nvert), wt = wave_sum(ntri);
    if (__lane_id() == 0) {
        atomicAdd(&sv, wv);
        atomicAdd(&st, wt);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        blk_v[blockIdx.x] = sv;
        blk_t[blockIdx.x] = st;
    }
}

// Pass 3a: vertex offsets of every voxel with cut edges, and the vertices.
__global__ __launch_bounds__(kMcBlock) void mc_vertex_kernel(McField F, float h, float ox, float oy, float oz,
                                                             const uint8_t* __restrict__ info,
                                                             const int64_t* __restrict__ vstart,
                                                             uint32_t* __restrict__ voff, float* __restrict__ verts) {
    __shared__ uint32_t sh[kMcBlock];
    const int64_t v = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
    const int mask = (v < F.n) ? (info[v] & 7) : 0;
    uint32_t total;
    const uint32_t excl = block_exclusive((uint32_t)__popc(mask), sh, &total);
    if (mask == 0) return;
    const int64_t base = vstart[blockIdx.x] + excl;
    voff[v] = (uint32_t)base;
    int x, y, z;
    mc_coords(F, v, x, y, z);
    const float v0 = mc_val(F, x, y, z) - F.iso;
    const float p[3] = {ox + (float)x * h, oy + (float)y * h, oz + (float)z * h};
    int r = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!(mask & (1 << a))) continue;
        const float v1 = mc_val(F, x + (a == 0), y + (a == 1), z + (a == 2)) - F.iso;
        const float den = v0 - v1;
        const float t = (fabsf(den) > 1e-7f) ? fminf(fmaxf(v0 / den, 0.001f), 0.999f) : 0.5f;
        float* o = verts + 3 * (base + r);
        o[0] = (a == 0) ? p[0] + t * h : p[0];
        o[1] = (a == 1) ? p[1] + t * h : p[1];
        o[2] = (a == 2) ? p[2] + t * h : p[2];
        ++r;
    }
}

// Pass 3b: triangles of every cube, (e0, e2, e1) per table triple.
__global__ __launch_bounds__(kMcBlock) void mc_triangle_kernel(McField F, const uint8_t* __restrict__ in,
                                                               const uint8_t* __restrict__ info,
                                                               const int64_t* __restrict__ tstart,
                                                               const uint32_t* __restrict__ voff,
                                                               uint32_t* __restrict__ tris) {
    __shared__ uint32_t sh[kMcBlock];
    const int64_t v = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
    const uint32_t ntri = (v < F.n) ? (uint32_t)(info[v] >> 3) : 0u;
    uint32_t total;
    const uint32_t excl = block_exclusive(ntri, sh, &total);
    if (ntri == 0) return;
    const int64_t ys = F.dx, zs = (int64_t)F.dx * F.dy;
    const int c = mc_case_bytes(in, v, ys, zs);
    int64_t out = tstart[blockIdx.x] + excl;
    for (int k = kMcTriStart[c]; k < kMcTriStart[c + 1]; k += 3, ++out) {
        uint32_t id[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int e = kMcTriEdges[k + j];
            const int oc = c_owner[e][0], a = c_owner[e][1];
            const int64_t o = v + c_corner[oc][0] + c_corner[oc][1] * ys + c_corner[oc][2] * zs;
            id[j] = voff[o] + (uint32_t)__popc(info[o] & 7 & ((1 << a) - 1));
        }
        tris[3 * out] = id[0];
        tris[3 * out + 1] = id[2];
        tris[3 * out + 2] = id[1];
    }
}

// Voxels with value <= 0, one atomic per wave.
__global__ __launch_bounds__(256) void interior_count_kernel(const float* __restrict__ g, int64_t n,
                                                             unsigned long long* __restrict__ count) {
    uint32_t c = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        c += (g[i] <= 0.0f) ? 1u : 0u;
    const uint32_t w = wave_sum(c);
    if (__lane_id() == 0 && w) atomicAdd(count, (unsigned long long)w);
}

}  // namespace

int64_t mc_blocks(const int32_t dims[3]) {
    const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    return (n + kMcBlock - 1) / kMcBlock;
}

hipError_t launch_mc_classify(const float* g, const int32_t dims[3], float h, float iso, uint8_t* in, uint8_t* info,
                              uint32_t* blk_v, uint32_t* blk_t, hipStream_t s) {
    const McField F{g, dims[0], dims[1], dims[2], iso, iso + h, (int64_t)dims[0] * dims[1] * dims[2]};
    const unsigned ib = (unsigned)std::min<int64_t>(mc_blocks(dims), 65536);
    mc_inside_kernel<<<ib, kMcBlock, 0, s>>>(F, in);
    mc_classify_kernel<<<(unsigned)mc_blocks(dims), kMcBlock, 0, s>>>(F, in, info, blk_v, blk_t);
    return hipGetLastError();
}

hipError_t launch_mc_emit(const float* g, const int32_t dims[3], float h, const float origin[3], float iso,
                          const uint8_t* in, const uint8_t* info, const int64_t* vstart, const int64_t* tstart,
                          uint32_t* voff, float* verts, uint32_t* tris, hipStream_t s) {
    const McField F{g, dims[0], dims[1], dims[2], iso, iso + h, (int64_t)dims[0] * dims[1] * dims[2]};
    const unsigned nb = (unsigned)mc_blocks(dims);
    mc_vertex_kernel<<<nb, kMcBlock, 0, s>>>(F, h, origin[0], origin[1], origin[2], info, vstart, voff, verts);
    mc_triangle_kernel<<<nb, kMcBlock, 0, s>>>(F, in, info, tstart, voff, tris);
    return hipGetLastError();
}

hipError_t launch_interior_count(const float* g, int64_t n, unsigned long long* count, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    interior_count_kernel<<<(unsigned)b, 256, 0, s>>>(g, n, count);
    return hipGetLastError();
}

}  // namespace rhz
