// k_fsm.hip -- [D] directional fast-sweeping passes and the standalone
// red-black eikonal.
//
// Schedule (bit-for-bit shared with the CPU oracle): one sweep walks every
// line of the swept axis sequentially (Gauss-Seidel along the line: the cell
// behind the walker is its new value), while every cross-line neighbour is read
// from the source buffer (the field before this sweep).  The result goes to a
// second buffer, so concurrent lines never observe each other's writes and the
// output is deterministic.  This is one legitimate interleaving of the
// reference gpu_eikonal_fsm_sweep_kernel (cuda_rasterizer.cu:433-513), whose
// Godunov update is not monotone in the largest neighbour and therefore
// schedule-dependent.
//
// Memory layout is x fastest.  y- and z-sweeps put 64 consecutive x in one
// wave, so every step is a coalesced 256-byte row access.  The x-sweep stages
// 64 lines x 32 columns (plus one halo row/column on each side) through LDS so
// global traffic stays row-coalesced while each lane walks its own line.
#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

// ---- y- and z-sweeps -------------------------------------------------------
// Block = 64 (x) x 4 (the other cross axis).  The walker keeps the current
// cell, the cell ahead (prefetched one step early) and the new value behind.
template <int AXIS>
__global__ __launch_bounds__(256) void sweep_yz_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                       int dx, int dy, int dz, int nbx, float h, int dir,
                                                       int* __restrict__ changed) {
    const int bx = blockIdx.x % nbx;
    const int bo = blockIdx.x / nbx;
    const int x = bx * 64 + threadIdx.x;
    const int o = bo * 4 + threadIdx.y;  // y for AXIS 2, z for AXIS 1
    const int olim = (AXIS == 2) ? dy : dz;
    const bool live = (x < dx) && (o < olim);
    const int64_t ys = dx, zs = (int64_t)dx * dy;
    const int len = (AXIS == 2) ? dz : dy;
    const int64_t as = (AXIS == 2) ? zs : ys;  // stride along the walk
    const int64_t os = (AXIS == 2) ? ys : zs;  // stride of the other cross axis
    bool chg = false;
    if (live) {
        const int64_t base = (AXIS == 2) ? ((int64_t)o * ys + x) : ((int64_t)o * zs + x);
        const bool xlo = x > 0, xhi = x < dx - 1, olo = o > 0, ohi = o < olim - 1;
        int q = (dir > 0) ? 0 : len - 1;
        int64_t i = base + (int64_t)q * as;
        float cur = src[i];
        float behind = kFar;
        for (int s = 0; s < len; ++s, q += dir, i += dir * as) {
            const bool more = s < len - 1;
            float ahead = more ? src[i + dir * as] : kFar;
            float xm = xlo ? src[i - 1] : kFar;
            float xp = xhi ? src[i + 1] : kFar;
            float om = olo ? src[i - os] : kFar;
            float op = ohi ? src[i + os] : kFar;
            float lo = (q > 0) ? ((dir > 0) ? behind : ahead) : kFar;
            float hi = (q < len - 1) ? ((dir > 0) ? ahead : behind) : kFar;
            float ax = fminf(xm, xp);
            float u;
            if (AXIS == 2) {
                u = godunov(ax, fminf(om, op), fminf(lo, hi), h);
            } else {
                u = godunov(ax, fminf(lo, hi), fminf(om, op), h);
            }
            float nv = cur;
            if (u < cur) {
                nv = u;
                chg = true;
            }
            dst[i] = nv;
            behind = nv;
            cur = ahead;
        }
    }
    if (__ballot(chg) != 0 && __lane_id() == 0) atomicOr(changed, 1);
}

// ---- x-sweep -----------------------------------------------------------------
constexpr int kXLines = 64;  // lines (y) per block, one wave
constexpr int kXChunk = 32;  // columns per LDS chunk

__global__ __launch_bounds__(kXLines) void sweep_x_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                          int dx, int dy, int dz, int nby, float h, int dir,
                                                          int* __restrict__ changed) {
    // T: rows y0-1 .. y0+64, columns xs-1 .. xs+kXChunk (old values)
    __shared__ float T[kXLines + 2][kXChunk + 3];
    __shared__ float N[kXLines][kXChunk + 1];
    const int lane = threadIdx.x;
    const int by = blockIdx.x % nby;
    const int z = blockIdx.x / nby;
    const int y0 = by * kXLines;
    const int y = y0 + lane;
    const bool live = y < dy;
    const int64_t ys = dx, zs = (int64_t)dx * dy;
    const int64_t plane = (int64_t)z * zs;
    const bool zlo = z > 0, zhi = z < dz - 1;
    const int nchunks = (dx + kXChunk - 1) / kXChunk;
    float behind = kFar;
    bool chg = false;
    for (int c = 0; c < nchunks; ++c) {
        const int ci = (dir > 0) ? c : (nchunks - 1 - c);
        const int xs = ci * kXChunk;
        // stage old values: 66 rows x (kXChunk + 2) columns
        constexpr int W = kXChunk + 2;
        for (int k = lane; k < (kXLines + 2) * W; k += kXLines) {
            int r = k / W, cc = k - r * W;
            int gy = y0 - 1 + r, gx = xs - 1 + cc;
            float v = kFar;
            if (gy >= 0 && gy < dy && gx >= 0 && gx < dx) v = src[plane + (int64_t)gy * ys + gx];
            T[r][cc] = v;
        }
        __syncthreads();
        if (live) {
            const int64_t row = plane + (int64_t)y * ys;
            for (int j = 0; j < kXChunk; ++j) {
                const int jj = (dir > 0) ? j : (kXChunk - 1 - j);
                const int gx = xs + jj;
                if (gx >= dx) continue;
                const int cc = jj + 1;
                float cur = T[lane + 1][cc];
                float lo = (gx > 0) ? ((dir > 0) ? behind : T[lane + 1][cc - 1]) : kFar;
                float hi = (gx < dx - 1) ? ((dir > 0) ? T[lane + 1][cc + 1] : behind) : kFar;
                float ym = (y > 0) ? T[lane][cc] : kFar;
                float yp = (y < dy - 1) ? T[lane + 2][cc] : kFar;
                float zm = zlo ? src[row - zs + gx] : kFar;
                float zp = zhi ? src[row + zs + gx] : kFar;
                float u = godunov(fminf(lo, hi), fminf(ym, yp), fminf(zm, zp), h);
                float nv = cur;
                if (u < cur) {
                    nv = u;
                    chg = true;
                }
                N[lane][jj] = nv;
                behind = nv;
            }
        }
        __syncthreads();
        for (int k = lane; k < kXLines * kXChunk; k += kXLines) {
            int r = k / kXChunk, cc = k - r * kXChunk;
            int gy = y0 + r, gx = xs + cc;
            if (gy < dy && gx < dx) dst[plane + (int64_t)gy * ys + gx] = N[r][cc];
        }
        __syncthreads();
    }
    if (__ballot(chg) != 0 && lane == 0) atomicOr(changed, 1);
}

// ---- standalone red-black eikonal (cuda_rasterizer.cu:1536-1578) ---------------
__global__ void redblack_kernel(float* __restrict__ g, int dx, int dy, int dz, float h, int parity) {
    int64_t n = (int64_t)dx * dy * dz;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t ys = dx, zs = (int64_t)dx * dy;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int x = (int)(i % dx), y = (int)((i / dx) % dy), z = (int)(i / zs);
        if (x <= 0 || x >= dx - 1 || y <= 0 || y >= dy - 1 || z <= 0 || z >= dz - 1) continue;
        if (((x + y + z) & 1) != parity) continue;
        float v = g[i];
        float current = fabsf(v);
        float sg = (v >= 0.0f) ? 1.0f : -1.0f;
        float ax = fminf(fabsf(g[i - 1]), fabsf(g[i + 1]));
        float ay = fminf(fabsf(g[i - ys]), fabsf(g[i + ys]));
        float az = fminf(fabsf(g[i - zs]), fabsf(g[i + zs]));
        float u = godunov(ax, ay, az, h);
        if (u < current) g[i] = u * sg;
    }
}

}  // namespace

hipError_t launch_sweep(const float* src, float* dst, const Grid& g, int axis, int dir, int* changed,
                        hipStream_t s) {
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    if (axis == 0) {
        int nby = (dy + kXLines - 1) / kXLines;
        int64_t blocks = (int64_t)nby * dz;
        sweep_x_kernel<<<(unsigned)blocks, kXLines, 0, s>>>(src, dst, dx, dy, dz, nby, g.h, dir, changed);
    } else {
        int nbx = (dx + 63) / 64;
        int other = (axis == 2) ? dy : dz;
        int64_t blocks = (int64_t)nbx * ((other + 3) / 4);
        dim3 blk(64, 4);
        if (axis == 2)
            sweep_yz_kernel<2><<<(unsigned)blocks, blk, 0, s>>>(src, dst, dx, dy, dz, nbx, g.h, dir, changed);
        else
            sweep_yz_kernel<1><<<(unsigned)blocks, blk, 0, s>>>(src, dst, dx, dy, dz, nbx, g.h, dir, changed);
    }
    return hipGetLastError();
}

hipError_t launch_redblack(float* g, const int32_t dims[3], float h, int parity, hipStream_t s) {
    int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    redblack_kernel<<<(unsigned)b, 256, 0, s>>>(g, dims[0], dims[1], dims[2], h, parity);
    return hipGetLastError();
}

}  // namespace rhz
