// k_splat.hip -- [A] exact +-band distance splat.
//
// One 256-thread block per triangle walks the triangle's band-dilated box
// (layer by layer in z, x fastest inside a layer, so neighbouring lanes hit
// neighbouring voxels) and folds the
// exact point-triangle distance into the grid with an unsigned atomic min on
// the float bit pattern (valid because every value is a non-negative float;
// the grid starts at the 1e30 sentinel).  The min over exact values does not
// depend on arrival order, so the band is bit-identical to a sequential
// restatement of reference gpu_splat_band_distances_kernel
// (cuda_rasterizer.cu:211-268): same box reject, same distance arithmetic,
// same voxel ranges (x/y by ceil/floor as :242-245, z by the per-layer test
// of :234).
#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int clamp_to_int(float v) {
    v = fminf(fmaxf(v, -2.0e9f), 2.0e9f);
    return (int)v;
}

__global__ __launch_bounds__(kBlock) void splat_kernel(const float* __restrict__ tris, int64_t nt,
                                                       int mnx, int mny, int mnz, int dx, int dy, int dz,
                                                       float h, float bw, unsigned* __restrict__ grid) {
    const int64_t t = blockIdx.x;
    if (t >= nt) return;
    const Tri9 tr = load_tri(tris, t);
    const float tx0 = fminf(tr.ax, fminf(tr.bx, tr.cx)), tx1 = fmaxf(tr.ax, fmaxf(tr.bx, tr.cx));
    const float ty0 = fminf(tr.ay, fminf(tr.by, tr.cy)), ty1 = fmaxf(tr.ay, fmaxf(tr.by, tr.cy));
    const float tz0 = fminf(tr.az, fminf(tr.bz, tr.cz)), tz1 = fmaxf(tr.az, fmaxf(tr.bz, tr.cz));
    const float band_sq = bw * bw;

    int x0 = clamp_to_int(ceilf((tx0 - bw) / h)) - mnx;
    int x1 = clamp_to_int(floorf((tx1 + bw) / h)) - mnx;
    int y0 = clamp_to_int(ceilf((ty0 - bw) / h)) - mny;
    int y1 = clamp_to_int(floorf((ty1 + bw) / h)) - mny;
    x0 = max(x0, 0);
    y0 = max(y0, 0);
    x1 = min(x1, dx - 1);
    y1 = min(y1, dy - 1);
    if (x1 < x0 || y1 < y0) return;

    // z layers kept by the reference test !(wz < tz0 - bw || wz > tz1 + bw);
    // wz grows with z, so the kept set is one interval found from an estimate
    // and exact predicate steps.
    const float zlo = tz0 - bw, zhi = tz1 + bw;
    int z0 = clamp_to_int(ceilf(zlo / h)) - mnz;
    z0 = min(max(z0, 0), dz);
    while (z0 > 0 && !(lattice(mnz, z0 - 1, h) < zlo)) --z0;
    while (z0 < dz && lattice(mnz, z0, h) < zlo) ++z0;
    int z1 = clamp_to_int(floorf(zhi / h)) - mnz;
    z1 = min(max(z1, -1), dz - 1);
    while (z1 < dz - 1 && !(lattice(mnz, z1 + 1, h) > zhi)) ++z1;
    while (z1 >= 0 && lattice(mnz, z1, h) > zhi) --z1;
    if (z1 < z0) return;

    // layer by layer in z; inside a layer the x-y box is indexed linearly
    // (k -> (k / nx, k % nx) by a float reciprocal and an exact correction)
    const int nx = x1 - x0 + 1, ny = y1 - y0 + 1, nxy = nx * ny;
    const float inv_nx = 1.0f / (float)nx;
    const int64_t zs = (int64_t)dx * dy;
    for (int z = z0; z <= z1; ++z) {
        const float pz = lattice(mnz, z, h);
        const float bdz = (pz < tz0) ? (tz0 - pz) : ((pz > tz1) ? (pz - tz1) : 0.0f);
        for (int k = (int)threadIdx.x; k < nxy; k += kBlock) {
            int ky = (int)((float)k * inv_nx);
            if (ky * nx > k) --ky;
            if ((ky + 1) * nx <= k) ++ky;
            const int x = x0 + (k - ky * nx), y = y0 + ky;
            const float px = lattice(mnx, x, h), py = lattice(mny, y, h);
            const float bdx = (px < tx0) ? (tx0 - px) : ((px > tx1) ? (px - tx1) : 0.0f);
            const float bdy = (py < ty0) ? (ty0 - py) : ((py > ty1) ? (py - ty1) : 0.0f);
            if (bdx * bdx + bdy * bdy + bdz * bdz > band_sq) continue;
            const float d2 = point_tri_dist2(px, py, pz, tr);
            if (d2 <= band_sq) {
                const float dist = sqrtf(d2);
                atomicMin(grid + (int64_t)z * zs + (int64_t)y * dx + x, __float_as_uint(dist));
            }
        }
    }
}

}  // namespace

hipError_t launch_splat(const float* tris, int64_t nt, const Grid& g, float bw, float* d, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    splat_kernel<<<(unsigned)nt, kBlock, 0, s>>>(tris, nt, g.mn[0], g.mn[1], g.mn[2], g.dims[0], g.dims[1],
                                                g.dims[2], g.h, bw, reinterpret_cast<unsigned*>(d));
    return hipGetLastError();
}

}  // namespace rhz
