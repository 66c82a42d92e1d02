// k_winding.hip -- exact generalised winding numbers (FP32 VALU-bound).
//
// launch_winding evaluates w(p) = sum_t Omega_t(p) / (4 pi) over ALL
// triangles for a list of lattice voxels.  Work is a 2-D grid: 256 points per
// block x `splits` triangle ranges, so a few thousand probe points still fill
// 256 CUs.  A lane owns one point; every lane of a wave reads the same
// triangle at the same time, so triangle loads are wave-uniform scalar loads
// (no LDS staging, no barriers).  Each lane keeps compensated (Kahan) FP32
// partials (two interleaved accumulators).  A finalize pass adds the
// partials in split order (deterministic), and every point whose ||w| - 0.5| is
// below fp64_band is queued for a binary64 recompute, which decides its sign.
//
// Precision 0 reproduces the reference arithmetic instead: plain FP32 sum in
// triangle order with the reference operation order (one split).
#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

constexpr int kWBlock = 256;
constexpr int kWTile = 256;
constexpr int64_t kMinTrisPerSplit = 1024;
constexpr int kTargetBlocks = 4096;

struct PointSrc {
    const int64_t* idx;
    int64_t base;
    int mnx, mny, mnz, dx, dy;
    float h;
    __device__ __forceinline__ int64_t voxel(int64_t i) const { return idx ? idx[i] : base + i; }
    __device__ __forceinline__ void pos(int64_t v, float& px, float& py, float& pz) const {
        int x = (int)(v % dx);
        int64_t r = v / dx;
        int y = (int)(r % dy);
        int z = (int)(r / dy);
        px = lattice(mnx, x, h);
        py = lattice(mny, y, h);
        pz = lattice(mnz, z, h);
    }
};

__device__ __forceinline__ void stage_tile(const float* __restrict__ tris, int64_t tb, int cnt, float* s) {
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        Tri9 t = load_tri(tris, tb + k);
        s[0 * kWTile + k] = t.ax;
        s[1 * kWTile + k] = t.ay;
        s[2 * kWTile + k] = t.az;
        s[3 * kWTile + k] = t.bx;
        s[4 * kWTile + k] = t.by;
        s[5 * kWTile + k] = t.bz;
        s[6 * kWTile + k] = t.cx;
        s[7 * kWTile + k] = t.cy;
        s[8 * kWTile + k] = t.cz;
    }
}

__device__ __forceinline__ Tri9 tile_tri(const float* s, int k) {
    return Tri9{s[k], s[kWTile + k], s[2 * kWTile + k], s[3 * kWTile + k], s[4 * kWTile + k],
                s[5 * kWTile + k], s[6 * kWTile + k], s[7 * kWTile + k], s[8 * kWTile + k]};
}

// Every lane owns one point; all lanes walk the same triangles, so triangle
// loads are wave-uniform (scalar loads, no LDS staging, no barriers).
template <int PREC>
__global__ __launch_bounds__(kWBlock) void winding_partial_kernel(const float* __restrict__ tris, int64_t nt,
                                                                  PointSrc src, int64_t n, int64_t per_split,
                                                                  float2* __restrict__ part) {
    const int64_t i = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.y * per_split;
    const int64_t t1 = min(nt, t0 + per_split);
    if (i >= n) return;
    float px, py, pz;
    src.pos(src.voxel(i), px, py, pz);
    float sum = 0.0f, comp = 0.0f;
    if (PREC == 0) {
        for (int64_t t = t0; t < t1; ++t) sum += solid_angle_ref(px, py, pz, load_tri(tris, t));
    } else {
        // two interleaved compensated accumulators (even / odd triangles)
        float sum2 = 0.0f, comp2 = 0.0f;
        int64_t t = t0;
        for (; t + 1 < t1; t += 2) {
            float a = solid_angle_fast(px, py, pz, load_tri(tris, t));
            float b = solid_angle_fast(px, py, pz, load_tri(tris, t + 1));
            float y = a - comp;
            float u = sum + y;
            comp = (u - sum) - y;
            sum = u;
            float y2 = b - comp2;
            float u2 = sum2 + y2;
            comp2 = (u2 - sum2) - y2;
            sum2 = u2;
        }
        if (t < t1) {
            float a = solid_angle_fast(px, py, pz, load_tri(tris, t));
            float y = a - comp;
            float u = sum + y;
            comp = (u - sum) - y;
            sum = u;
        }
        // fold the second accumulator into the first
        float terms[2] = {sum2, -comp2};
        for (int j = 0; j < 2; ++j) {
            float y = terms[j] - comp;
            float u = sum + y;
            comp = (u - sum) - y;
            sum = u;
        }
    }
    part[(int64_t)blockIdx.y * n + i] = make_float2(sum, comp);
}

__global__ void winding_finalize_kernel(const float2* __restrict__ part, int nsplit, int64_t n, int prec,
                                        float band, double* __restrict__ w, int64_t* __restrict__ fp64_list,
                                        unsigned long long* fp64_cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool want = false;
    if (i < n) {
        float wf;
        if (prec == 0) {
            float s = 0.0f;
            for (int k = 0; k < nsplit; ++k) s += part[(int64_t)k * n + i].x;
            wf = s * kInv4Pi;
        } else {
            float s = 0.0f, c = 0.0f;
            for (int k = 0; k < nsplit; ++k) {
                float2 p = part[(int64_t)k * n + i];
                float terms[2] = {p.x, -p.y};
                for (int j = 0; j < 2; ++j) {
                    float y = terms[j] - c;
                    float t = s + y;
                    c = (t - s) - y;
                    s = t;
                }
            }
            wf = (s - c) * kInv4Pi;
            want = fabsf(fabsf(wf) - 0.5f) < band;
        }
        w[i] = (double)wf;
    }
    if (fp64_list) {
        long long slot = wave_append(fp64_cnt, want);
        if (want) fp64_list[slot] = i;
    }
}

__global__ __launch_bounds__(kWBlock) void winding_fp64_kernel(const float* __restrict__ tris, int64_t nt,
                                                               PointSrc src, const int64_t* __restrict__ list,
                                                               int64_t m, double* __restrict__ w) {
    const int64_t j = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    if (j >= m) return;
    const int64_t i = list[j];
    float fx, fy, fz;
    src.pos(src.voxel(i), fx, fy, fz);
    const double px = fx, py = fy, pz = fz;
    double sum = 0.0;
    for (int64_t t = 0; t < nt; ++t) sum += solid_angle_f64(px, py, pz, load_tri(tris, t));
    w[i] = sum / (4.0 * 3.14159265358979323846);
}

// The same recompute with the list length read on the device (no host round
// trip): the grid covers the list's capacity and lanes past the count exit.
__global__ __launch_bounds__(kWBlock) void winding_fp64_count_kernel(const float* __restrict__ tris, int64_t nt,
                                                                     PointSrc src, const int64_t* __restrict__ list,
                                                                     const unsigned long long* __restrict__ count,
                                                                     double* __restrict__ w) {
    const int64_t j = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    if ((unsigned long long)j >= *count) return;
    const int64_t i = list[j];
    float fx, fy, fz;
    src.pos(src.voxel(i), fx, fy, fz);
    const double px = fx, py = fy, pz = fz;
    double sum = 0.0;
    for (int64_t t = 0; t < nt; ++t) sum += solid_angle_f64(px, py, pz, load_tri(tris, t));
    w[i] = sum / (4.0 * 3.14159265358979323846);
}

// ---- legacy exact kernel -------------------------------------------------------
constexpr int kLegacyTile = 128;  // RASTER_TILE_SIZE (core/math_utils.cuh:35)

__global__ void legacy_tile_bbox_kernel(const float* __restrict__ tris, int64_t nt, float* __restrict__ bb) {
    int64_t tile = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t ntiles = (nt + kLegacyTile - 1) / kLegacyTile;
    if (tile >= ntiles) return;
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    int64_t s0 = tile * kLegacyTile, s1 = min(nt, s0 + kLegacyTile);
    for (int64_t k = s0; k < s1; ++k) {
        Tri9 t = load_tri(tris, k);
        lo[0] = fminf(lo[0], fminf(t.ax, fminf(t.bx, t.cx)));
        lo[1] = fminf(lo[1], fminf(t.ay, fminf(t.by, t.cy)));
        lo[2] = fminf(lo[2], fminf(t.az, fminf(t.bz, t.cz)));
        hi[0] = fmaxf(hi[0], fmaxf(t.ax, fmaxf(t.bx, t.cx)));
        hi[1] = fmaxf(hi[1], fmaxf(t.ay, fmaxf(t.by, t.cy)));
        hi[2] = fmaxf(hi[2], fmaxf(t.az, fmaxf(t.bz, t.cz)));
    }
    for (int j = 0; j < 3; ++j) {
        bb[6 * tile + j] = lo[j];
        bb[6 * tile + 3 + j] = hi[j];
    }
}

// Reference mesh_to_sdf_kernel (cuda_rasterizer.cu:97-185): per voxel, min
// distance over triangles with the per-tile AABB cull (:140-145), and the
// plain FP32 solid-angle sum over every triangle in order (:164-174).
__global__ __launch_bounds__(kWBlock) void legacy_kernel(const float* __restrict__ tris, int64_t nt,
                                                         const float* __restrict__ bb, PointSrc src, int64_t n,
                                                         float* __restrict__ out) {
    __shared__ float s[9 * kWTile];
    const int64_t i = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    const bool live = i < n;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (live) src.pos(i, px, py, pz);
    float mind = 1e12f, sum = 0.0f;
    const int64_t ntiles = (nt + kLegacyTile - 1) / kLegacyTile;
    for (int64_t tile = 0; tile < ntiles; ++tile) {
        const int64_t tb = tile * kLegacyTile;
        const int cnt = (int)min((int64_t)kLegacyTile, nt - tb);
        __syncthreads();
        stage_tile(tris, tb, cnt, s);
        __syncthreads();
        if (live) {
            const float* b = bb + 6 * tile;
            float bdx = (px < b[0]) ? (b[0] - px) : ((px > b[3]) ? (px - b[3]) : 0.0f);
            float bdy = (py < b[1]) ? (b[1] - py) : ((py > b[4]) ? (py - b[4]) : 0.0f);
            float bdz = (pz < b[2]) ? (b[2] - pz) : ((pz > b[5]) ? (pz - b[5]) : 0.0f);
            const bool can = (bdx * bdx + bdy * bdy + bdz * bdz) < mind;
            for (int k = 0; k < cnt; ++k) {
                Tri9 t = tile_tri(s, k);
                if (can) {
                    float d2 = point_tri_dist2(px, py, pz, t);
                    if (d2 < mind) mind = d2;
                }
                sum += solid_angle_ref(px, py, pz, t);
            }
        }
    }
    if (live) {
        float wn = sum * kInv4Pi;
        float sg = (fabsf(wn) >= 0.5f) ? -1.0f : 1.0f;
        out[i] = sg * sqrtf(mind);
    }
}

// Stages triangles tb .. tb + cnt - 1 of an indexed mesh into the LDS tile s
// (component-major: s[c * kWTile + k], c = a.x a.y a.z b.x ... c.z).
__device__ __forceinline__ void stage_indexed_tris(float* s, const float* __restrict__ verts,
                                                   const uint32_t* __restrict__ idx, int64_t tb, int cnt) {
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        const int64_t t = tb + k;
        const uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
        s[0 * kWTile + k] = verts[3 * (int64_t)i0];
        s[1 * kWTile + k] = verts[3 * (int64_t)i0 + 1];
        s[2 * kWTile + k] = verts[3 * (int64_t)i0 + 2];
        s[3 * kWTile + k] = verts[3 * (int64_t)i1];
        s[4 * kWTile + k] = verts[3 * (int64_t)i1 + 1];
        s[5 * kWTile + k] = verts[3 * (int64_t)i1 + 2];
        s[6 * kWTile + k] = verts[3 * (int64_t)i2];
        s[7 * kWTile + k] = verts[3 * (int64_t)i2 + 1];
        s[8 * kWTile + k] = verts[3 * (int64_t)i2 + 2];
    }
}

// ---- standalone winding numbers over an indexed mesh ------------------------
// Reference gpu_compute_winding_numbers_kernel (engine/mesh/cuda_mesh_ops.cu:
// 774-823): guard < 1e-7 skips the triangle, plain FP32 sum in index order.
__global__ __launch_bounds__(kWBlock) void winding_points_kernel(const float* __restrict__ verts,
                                                                 const uint32_t* __restrict__ idx, int64_t nt,
                                                                 const float* __restrict__ pts, int64_t np,
                                                                 float* __restrict__ out) {
    __shared__ float s[9 * kWTile];
    const int64_t i = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    const bool live = i < np;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (live) {
        px = pts[3 * i];
        py = pts[3 * i + 1];
        pz = pts[3 * i + 2];
    }
    float total = 0.0f;
    for (int64_t tb = 0; tb < nt; tb += kWTile) {
        const int cnt = (int)min((int64_t)kWTile, nt - tb);
        __syncthreads();
        stage_indexed_tris(s, verts, idx, tb, cnt);
        __syncthreads();
        if (live) {
            for (int k = 0; k < cnt; ++k) {
                Tri9 t = tile_tri(s, k);
                float ax = t.ax - px, ay = t.ay - py, az = t.az - pz;
                float bx = t.bx - px, by = t.by - py, bz = t.bz - pz;
                float cx = t.cx - px, cy = t.cy - py, cz = t.cz - pz;
                float la = sqrtf(ax * ax + ay * ay + az * az);
                float lb = sqrtf(bx * bx + by * by + bz * bz);
                float lc = sqrtf(cx * cx + cy * cy + cz * cz);
                if (la < 1e-7f || lb < 1e-7f || lc < 1e-7f) continue;
                float det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
                float dab = ax * bx + ay * by + az * bz;
                float dbc = bx * cx + by * cy + bz * cz;
                float dca = cx * ax + cy * ay + cz * az;
                float den = la * lb * lc + dab * lc + dbc * la + dca * lb;
                total += 2.0f * atan2f(det, den);
            }
        }
    }
    if (live) out[i] = total * kInv4Pi;
}

// ---- standalone point-to-mesh distances -------------------------------------
// Reference gpu_query_point_distances_kernel (engine/mesh/cuda_mesh_ops.cu:
// 668-703): the square root of the smallest squared point-triangle distance
// over every triangle of the indexed mesh (degenerate ones included; a NaN
// distance never replaces the running minimum), starting from 1e18.
__global__ __launch_bounds__(kWBlock) void point_distance_kernel(const float* __restrict__ verts,
                                                                 const uint32_t* __restrict__ idx, int64_t nt,
                                                                 const float* __restrict__ pts, int64_t np,
                                                                 float* __restrict__ out) {
    __shared__ float s[9 * kWTile];
    const int64_t i = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    const bool live = i < np;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (live) {
        px = pts[3 * i];
        py = pts[3 * i + 1];
        pz = pts[3 * i + 2];
    }
    float best = 1e18f;
    for (int64_t tb = 0; tb < nt; tb += kWTile) {
        const int cnt = (int)min((int64_t)kWTile, nt - tb);
        __syncthreads();
        stage_indexed_tris(s, verts, idx, tb, cnt);
        __syncthreads();
        if (live) {
            for (int k = 0; k < cnt; ++k) {
                const float d2 = point_tri_dist2(px, py, pz, tile_tri(s, k));
                if (d2 < best) best = d2;
            }
        }
    }
    if (live) out[i] = sqrtf(best);
}

// ---- binary64 distance to the nearest triangle (accuracy yardstick) -------------
// d2[i] = min over all triangles of the binary64 squared distance from voxel
// idx[i]'s lattice point (binary32 coordinates, as the field uses).  Blocks
// split the triangles into ranges; each lane keeps its range's minimum and
// merges it with an integer atomic min on the bit pattern (non-negative
// binary64 values order like their bit patterns), so the result does not
// depend on the order of the ranges.
__global__ __launch_bounds__(kWBlock) void distance_f64_kernel(const float* __restrict__ tris, int64_t nt,
                                                               PointSrc src, int64_t n, int64_t per_split,
                                                               unsigned long long* __restrict__ d2) {
    __shared__ float s[9 * kWTile];
    const int64_t i = (int64_t)blockIdx.x * kWBlock + threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.y * per_split;
    const int64_t t1 = min(nt, t0 + per_split);
    const bool live = i < n;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (live) {
        float fx, fy, fz;
        src.pos(src.voxel(i), fx, fy, fz);
        px = fx;
        py = fy;
        pz = fz;
    }
    double best = __longlong_as_double(0x7ff0000000000000ll);  // +inf
    for (int64_t tb = t0; tb < t1; tb += kWTile) {
        const int cnt = (int)min((int64_t)kWTile, t1 - tb);
        __syncthreads();
        stage_tile(tris, tb, cnt, s);
        __syncthreads();
        if (live) {
            for (int k = 0; k < cnt; ++k) best = fmin(best, point_tri_dist2_f64(px, py, pz, tile_tri(s, k)));
        }
    }
    if (live) atomicMin(d2 + i, (unsigned long long)__double_as_longlong(best));
}

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline int split_count(int64_t n, int64_t nt) {
    int64_t pblocks = (n + kWBlock - 1) / kWBlock;
    int64_t want = (kTargetBlocks + pblocks - 1) / pblocks;
    int64_t maxs = (nt + kMinTrisPerSplit - 1) / kMinTrisPerSplit;
    if (want > maxs) want = maxs;
    if (want < 1) want = 1;
    if (want > 65535) want = 65535;
    return (int)want;
}

PointSrc make_src(const Grid& g, const int64_t* idx, int64_t base) {
    PointSrc p;
    p.idx = idx;
    p.base = base;
    p.mnx = g.mn[0];
    p.mny = g.mn[1];
    p.mnz = g.mn[2];
    p.dx = g.dims[0];
    p.dy = g.dims[1];
    p.h = g.h;
    return p;
}

}  // namespace

size_t winding_scratch_bytes(int64_t n_points, int64_t n_tris) {
    return (size_t)split_count(n_points, n_tris) * (size_t)n_points * sizeof(float2) + 256;
}

hipError_t launch_winding(const float* tris, int64_t nt, const Grid& g, const int64_t* idx, int64_t base,
                          int64_t n, int precision, float fp64_band, double* w, void* scratch,
                          int64_t* fp64_list, unsigned long long* fp64_cnt, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    PointSrc src = make_src(g, idx, base);
    int nsplit = (precision == 0) ? 1 : split_count(n, nt);
    int64_t per_split = (nt + nsplit - 1) / nsplit;
    // keep split boundaries on tile multiples
    per_split = ((per_split + kWTile - 1) / kWTile) * kWTile;
    nsplit = (int)((nt + per_split - 1) / per_split);
    if (nsplit < 1) nsplit = 1;
    float2* part = reinterpret_cast<float2*>(scratch);
    dim3 grid((unsigned)((n + kWBlock - 1) / kWBlock), (unsigned)nsplit);
    if (precision == 0)
        winding_partial_kernel<0><<<grid, kWBlock, 0, s>>>(tris, nt, src, n, per_split, part);
    else
        winding_partial_kernel<1><<<grid, kWBlock, 0, s>>>(tris, nt, src, n, per_split, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    winding_finalize_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(
        part, nsplit, n, precision, fp64_band, w, precision == 0 ? nullptr : fp64_list, fp64_cnt);
    return hipGetLastError();
}

hipError_t launch_winding_fp64(const float* tris, int64_t nt, const Grid& g, const int64_t* idx, int64_t base,
                               const int64_t* fp64_list, int64_t n_fp64, double* w, hipStream_t s) {
    if (n_fp64 <= 0) return hipSuccess;
    PointSrc src = make_src(g, idx, base);
    winding_fp64_kernel<<<(unsigned)((n_fp64 + kWBlock - 1) / kWBlock), kWBlock, 0, s>>>(tris, nt, src, fp64_list,
                                                                                        n_fp64, w);
    return hipGetLastError();
}

hipError_t launch_winding_fp64_count(const float* tris, int64_t nt, const Grid& g, const int64_t* idx, int64_t base,
                                     const int64_t* fp64_list, const unsigned long long* count, int64_t cap, double* w,
                                     hipStream_t s) {
    if (cap <= 0) return hipSuccess;
    PointSrc src = make_src(g, idx, base);
    winding_fp64_count_kernel<<<(unsigned)((cap + kWBlock - 1) / kWBlock), kWBlock, 0, s>>>(tris, nt, src, fp64_list,
                                                                                           count, w);
    return hipGetLastError();
}

size_t legacy_scratch_bytes(int64_t nt) { return (size_t)((nt + kLegacyTile - 1) / kLegacyTile) * 6 * sizeof(float) + 256; }

hipError_t launch_legacy(const float* tris, int64_t nt, const Grid& g, float* out, void* scratch, hipStream_t s) {
    int64_t n = g.n();
    if (n <= 0) return hipSuccess;
    float* bb = reinterpret_cast<float*>(scratch);
    int64_t ntiles = (nt + kLegacyTile - 1) / kLegacyTile;
    if (ntiles > 0) legacy_tile_bbox_kernel<<<(unsigned)((ntiles + 255) / 256), 256, 0, s>>>(tris, nt, bb);
    PointSrc src = make_src(g, nullptr, 0);
    legacy_kernel<<<(unsigned)((n + kWBlock - 1) / kWBlock), kWBlock, 0, s>>>(tris, nt, bb, src, n, out);
    return hipGetLastError();
}

hipError_t launch_distance_f64(const float* tris, int64_t nt, const Grid& g, const int64_t* idx, int64_t n,
                               double* d2, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    unsigned long long* u = reinterpret_cast<unsigned long long*>(d2);
    fill_u64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(u, n, 0x7ff0000000000000ull);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nt <= 0) return e;
    PointSrc src = make_src(g, idx, 0);
    int nsplit = split_count(n, nt);
    int64_t per_split = (nt + nsplit - 1) / nsplit;
    per_split = ((per_split + kWTile - 1) / kWTile) * kWTile;
    nsplit = (int)((nt + per_split - 1) / per_split);
    dim3 grid((unsigned)((n + kWBlock - 1) / kWBlock), (unsigned)nsplit);
    distance_f64_kernel<<<grid, kWBlock, 0, s>>>(tris, nt, src, n, per_split, u);
    return hipGetLastError();
}

hipError_t launch_winding_points(const float* verts, const uint32_t* idx, int64_t nt, const float* pts, int64_t np,
                                 float* out, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    winding_points_kernel<<<(unsigned)((np + kWBlock - 1) / kWBlock), kWBlock, 0, s>>>(verts, idx, nt, pts, np, out);
    return hipGetLastError();
}

hipError_t launch_point_distances(const float* verts, const uint32_t* idx, int64_t nt, const float* pts, int64_t np,
                                  float* out, hipStream_t s) {
    if (np <= 0) return hipSuccess;
    point_distance_kernel<<<(unsigned)((np + kWBlock - 1) / kWBlock), kWBlock, 0, s>>>(verts, idx, nt, pts, np, out);
    return hipGetLastError();
}

}  // namespace rhz
