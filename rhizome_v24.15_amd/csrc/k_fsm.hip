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
// wave (every step is a coalesced 256-byte row access) and prefetch the
// inputs of the next kStep steps before walking them, so each wave keeps
// several rows in flight.  The x-sweep stages a 64-line x 32-column tile of
// the plane (with one halo row/column on each side) and the matching tiles of
// the planes above and below through LDS, so global traffic stays
// row-coalesced while each lane walks its own line.  Block order is remapped
// so neighbouring tiles share an XCD's L2.
#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

constexpr int kStep = 4;     // walk steps whose loads are issued together
constexpr int kYZRows = 8;   // waves per block in the y/z sweeps

struct SweepConst {
    float h, h2x2, h2x3, h057, hbig;
};

__device__ __forceinline__ SweepConst sweep_const(float h) {
    SweepConst c;
    c.h = h;
    c.h2x2 = 2.0f * h * h;
    c.h2x3 = 3.0f * h * h;
    c.h057 = 0.57f * h;
    c.hbig = 4096.0f * h;
    return c;
}

// ---- generic y- and z-sweeps (any shape) -------------------------------------
template <int AXIS>
__global__ __launch_bounds__(64 * kYZRows) void sweep_yz_kernel(const float* __restrict__ src,
                                                                 float* __restrict__ dst, int dx, int dy, int dz,
                                                                 int nbx, unsigned nblocks, float h, int dir,
                                                                 int* __restrict__ changed) {
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const int bx = (int)(lb % (unsigned)nbx);
    const int bo = (int)(lb / (unsigned)nbx);
    const int x = bx * 64 + threadIdx.x;
    const int o = bo * kYZRows + threadIdx.y;  // y for AXIS 2, z for AXIS 1
    const int olim = (AXIS == 2) ? dy : dz;
    const bool live = (x < dx) && (o < olim);
    const int64_t ys = dx, zs = (int64_t)dx * dy;
    const int len = (AXIS == 2) ? dz : dy;
    const int64_t as = (AXIS == 2) ? zs : ys;  // stride along the walk
    const int64_t os = (AXIS == 2) ? ys : zs;  // stride of the other cross axis
    const SweepConst K = sweep_const(h);
    bool chg = false;
    if (live) {
        const int64_t base = (AXIS == 2) ? ((int64_t)o * ys + x) : ((int64_t)o * zs + x);
        // neighbour offsets clamped onto the cell itself at the grid faces
        // (the loaded value is then replaced by the sentinel)
        const bool xlo = x > 0, xhi = x < dx - 1, olo = o > 0, ohi = o < olim - 1;
        const int64_t oxm = xlo ? -1 : 0, oxp = xhi ? 1 : 0, oom = olo ? -os : 0, oop = ohi ? os : 0;
        const int64_t step = (int64_t)dir * as;
        int q = (dir > 0) ? 0 : len - 1;
        int64_t i = base + (int64_t)q * as;
        float behind = kFar;
        float cur = src[i];
        for (int s0 = 0; s0 < len; s0 += kStep) {
            float c[kStep + 1], xm[kStep], xp[kStep], om[kStep], op[kStep];
            c[0] = cur;
#pragma unroll
            for (int k = 0; k < kStep; ++k) {
                const bool ok = s0 + k < len;
                const int64_t ik = ok ? i + k * step : i;  // clamped inside the line
                xm[k] = src[ik + oxm];
                xp[k] = src[ik + oxp];
                om[k] = src[ik + oom];
                op[k] = src[ik + oop];
                const bool okn = s0 + k + 1 < len;
                c[k + 1] = src[okn ? ik + step : ik];
            }
#pragma unroll
            for (int k = 0; k < kStep; ++k) {
                if (s0 + k >= len) break;
                const float ahead = c[k + 1];
                const float lo = (q > 0) ? ((dir > 0) ? behind : ahead) : kFar;
                const float hi = (q < len - 1) ? ((dir > 0) ? ahead : behind) : kFar;
                const float ax = fminf(xlo ? xm[k] : kFar, xhi ? xp[k] : kFar);
                const float ao = fminf(olo ? om[k] : kFar, ohi ? op[k] : kFar);
                const float al = fminf(lo, hi);
                float nv;
                if (AXIS == 2)
                    nv = godunov_min(c[k], ax, ao, al, K.h, K.h2x2, K.h2x3, K.h057, K.hbig);
                else
                    nv = godunov_min(c[k], ax, al, ao, K.h, K.h2x2, K.h2x3, K.h057, K.hbig);
                chg |= nv < c[k];
                dst[i] = nv;
                behind = nv;
                i += step;
                q += dir;
            }
            cur = c[kStep];
        }
    }
    if (__ballot(chg) != 0 && __lane_id() == 0) atomicOr(changed, 1);
}

// ---- generic x-sweep (any shape) ----------------------------------------------
constexpr int kXLines = 64;  // lines (y) per block, one wave
constexpr int kXChunk = 32;  // columns per LDS chunk

__global__ __launch_bounds__(kXLines) void sweep_x_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                          int dx, int dy, int dz, int nby, unsigned nblocks, float h,
                                                          int dir, int* __restrict__ changed) {
    // T: rows y0-1 .. y0+64, columns xs-1 .. xs+kXChunk (old values);
    // Z0/Z1: rows y0 .. y0+63 of planes z-1 / z+1; N: new values.
    __shared__ float T[kXLines + 2][kXChunk + 3];
    __shared__ float Z0[kXLines][kXChunk + 1];
    __shared__ float Z1[kXLines][kXChunk + 1];
    __shared__ float N[kXLines][kXChunk + 1];
    const int lane = threadIdx.x;
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const int by = (int)(lb % (unsigned)nby);
    const int z = (int)(lb / (unsigned)nby);
    const int y0 = by * kXLines;
    const int y = y0 + lane;
    const bool live = y < dy;
    const int64_t ys = dx, zs = (int64_t)dx * dy;
    const int64_t plane = (int64_t)z * zs;
    const bool zlo = z > 0, zhi = z < dz - 1;
    const int nchunks = (dx + kXChunk - 1) / kXChunk;
    const SweepConst K = sweep_const(h);
    float behind = kFar;
    bool chg = false;
    for (int c = 0; c < nchunks; ++c) {
        const int ci = (dir > 0) ? c : (nchunks - 1 - c);
        const int xs = ci * kXChunk;
        constexpr int W = kXChunk + 2;
        for (int k = lane; k < (kXLines + 2) * W; k += kXLines) {
            const int r = k / W, cc = k - r * W;
            const int gy = y0 - 1 + r, gx = xs - 1 + cc;
            float v = kFar;
            if (gy >= 0 && gy < dy && gx >= 0 && gx < dx) v = src[plane + (int64_t)gy * ys + gx];
            T[r][cc] = v;
        }
        for (int k = lane; k < kXLines * kXChunk; k += kXLines) {
            const int r = k / kXChunk, cc = k - r * kXChunk;
            const int gy = y0 + r, gx = xs + cc;
            const bool in = gy < dy && gx < dx;
            const int64_t g = plane + (int64_t)gy * ys + gx;
            Z0[r][cc] = (in && zlo) ? src[g - zs] : kFar;
            Z1[r][cc] = (in && zhi) ? src[g + zs] : kFar;
        }
        __syncthreads();
        if (live) {
            for (int j = 0; j < kXChunk; ++j) {
                const int jj = (dir > 0) ? j : (kXChunk - 1 - j);
                const int gx = xs + jj;
                if (gx >= dx) continue;
                const int cc = jj + 1;
                const float cur = T[lane + 1][cc];
                const float lo = (gx > 0) ? ((dir > 0) ? behind : T[lane + 1][cc - 1]) : kFar;
                const float hi = (gx < dx - 1) ? ((dir > 0) ? T[lane + 1][cc + 1] : behind) : kFar;
                const float ym = (y > 0) ? T[lane][cc] : kFar;
                const float yp = (y < dy - 1) ? T[lane + 2][cc] : kFar;
                const float nv = godunov_min(cur, fminf(lo, hi), fminf(ym, yp), fminf(Z0[lane][jj], Z1[lane][jj]),
                                             K.h, K.h2x2, K.h2x3, K.h057, K.hbig);
                chg |= nv < cur;
                N[lane][jj] = nv;
                behind = nv;
            }
        }
        __syncthreads();
        for (int k = lane; k < kXLines * kXChunk; k += kXLines) {
            const int r = k / kXChunk, cc = k - r * kXChunk;
            const int gy = y0 + r, gx = xs + cc;
            if (gy < dy && gx < dx) dst[plane + (int64_t)gy * ys + gx] = N[r][cc];
        }
        __syncthreads();
    }
    if (__ballot(chg) != 0 && lane == 0) atomicOr(changed, 1);
}

// ---- lane-walker sweep (the fast path) ------------------------------------------
// One wave = 64 consecutive cells of the fast (contiguous) axis F; every lane
// walks its line along axis W; C is the remaining cross axis.  Each axis
// feeds one Godunov argument slot (0 = x, 1 = y, 2 = z) so the same kernel
// serves the y- and z-sweeps on the x-fastest layout and the x-sweep on a
// y-fastest copy.  The wave's row pointer is uniform (scalar registers);
// lanes add 32-bit offsets.  A missing neighbour at a grid face reads the
// opposite neighbour instead: fminf(v, v) = fminf(1e30, v) for every stored
// value (all are <= 1e30), so no per-step select is needed (F and C must have
// at least 2 cells).
template <int SF, int SW, int SC, int DIR>
__global__ __launch_bounds__(64 * kYZRows) void sweep_lane_kernel(const float* __restrict__ src,
                                                                   float* __restrict__ dst, int nf, int nw, int nc,
                                                                   int64_t ws, int64_t cs, int nbf, unsigned nblocks,
                                                                   float h, int* __restrict__ changed) {
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const int bf = (int)(lb % (unsigned)nbf);
    const int bc = (int)(lb / (unsigned)nbf);
    const int c = __builtin_amdgcn_readfirstlane(bc * kYZRows + (int)threadIdx.y);
    if (c >= nc) return;
    const int lane = threadIdx.x;
    const int f = bf * 64 + lane;
    const bool live = f < nf;
    const int fl = live ? f : nf - 1;
    const int ofm = (fl > 0) ? -1 : 1;
    const int ofp = (fl < nf - 1) ? 1 : -1;
    const int64_t ocm = (c > 0) ? -cs : cs;
    const int64_t ocp = (c < nc - 1) ? cs : -cs;
    const int om = (int)ocm, op = (int)ocp;  // |cs| < 2^29 (checked on the host)
    const SweepConst K = sweep_const(h);
    const int64_t wave_base = (int64_t)c * cs;  // lane offsets fl carry the fast coordinate
    const int64_t step = (int64_t)DIR * ws;
    int64_t row = wave_base + (int64_t)((DIR > 0) ? 0 : nw - 1) * ws;
    float behind = kFar;
    float cur = src[row + fl];
    bool chg = false;
    for (int s0 = 0; s0 < nw; s0 += kStep) {
        float cc[kStep + 1], fm[kStep], fp[kStep], cm[kStep], cp[kStep];
        cc[0] = cur;
#pragma unroll
        for (int k = 0; k < kStep; ++k) {
            const int64_t rk = (s0 + k < nw) ? row + k * step : row;
            const float* __restrict__ rp = src + rk;
            fm[k] = rp[fl + ofm];
            fp[k] = rp[fl + ofp];
            cm[k] = rp[fl + om];
            cp[k] = rp[fl + op];
            cc[k + 1] = (s0 + k + 1 < nw) ? src[rk + step + fl] : kFar;
        }
#pragma unroll
        for (int k = 0; k < kStep; ++k) {
            if (s0 + k >= nw) break;
            const float a_f = fminf(fm[k], fp[k]);
            const float a_c = fminf(cm[k], cp[k]);
            const float a_w = fminf(behind, cc[k + 1]);
            float g[3];
            g[SF] = a_f;
            g[SW] = a_w;
            g[SC] = a_c;
            // The Godunov solve runs only when some lane of the wave can improve:
            // a uniform branch, so converged stretches (most of the far field
            // after the first round) cost a handful of instructions per cell.
            // Lanes that skip keep cc[k], which is what godunov_min's own exit
            // returns for them: bit-identical either way.
            const float lo = fminf(g[0], fminf(g[1], g[2]));
            const float hi = fmaxf(g[0], fmaxf(g[1], g[2]));
            const bool need = (cc[k] > lo + K.h057) || !(hi < K.hbig);
            float nv = cc[k];
            if (__ballot(need) != 0)
                nv = godunov_min(cc[k], g[0], g[1], g[2], K.h, K.h2x2, K.h2x3, K.h057, K.hbig);
            chg |= nv < cc[k];
            if (live) dst[row + fl] = nv;
            behind = nv;
            row += step;
        }
        cur = cc[kStep];
    }
    if (__ballot(chg) != 0 && lane == 0) atomicOr(changed, 1);
}

// ---- plane transpose (x-fastest <-> y-fastest) ----------------------------------
// Each z-plane is an (rows x cols) matrix; out[col][row] = in[row][col].
// 64x64 tiles through LDS (one padding column: conflict-free column reads).
__global__ __launch_bounds__(256) void transpose_planes_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                               int rows, int cols, int tiles_r, int tiles_c,
                                                               unsigned nblocks) {
    __shared__ float tile[64][65];
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const int per_plane = tiles_r * tiles_c;
    const int z = (int)(lb / (unsigned)per_plane);
    const int t = (int)(lb % (unsigned)per_plane);
    const int tr = t / tiles_c, tc = t % tiles_c;
    const int64_t plane = (int64_t)z * rows * cols;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    const int c0 = tc * 64, r0 = tr * 64;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int r = r0 + ty + 4 * k, c = c0 + tx;
        if (r < rows && c < cols) tile[ty + 4 * k][tx] = in[plane + (int64_t)r * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int c = c0 + ty + 4 * k, r = r0 + tx;
        if (r < rows && c < cols) out[plane + (int64_t)c * rows + r] = tile[tx][ty + 4 * k];
    }
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

namespace {

template <int SF, int SW, int SC>
hipError_t launch_lane(const float* src, float* dst, int nf, int nw, int nc, int64_t ws, int64_t cs, float h, int dir,
                       int* changed, hipStream_t s) {
    const int nbf = (nf + 63) / 64;
    const int64_t blocks = (int64_t)nbf * ((nc + kYZRows - 1) / kYZRows);
    dim3 blk(64, kYZRows);
    if (dir > 0)
        sweep_lane_kernel<SF, SW, SC, 1><<<(unsigned)blocks, blk, 0, s>>>(src, dst, nf, nw, nc, ws, cs, nbf,
                                                                         (unsigned)blocks, h, changed);
    else
        sweep_lane_kernel<SF, SW, SC, -1><<<(unsigned)blocks, blk, 0, s>>>(src, dst, nf, nw, nc, ws, cs, nbf,
                                                                          (unsigned)blocks, h, changed);
    return hipGetLastError();
}

hipError_t launch_generic_sweep(const float* src, float* dst, const Grid& g, int axis, int dir, int* changed,
                                hipStream_t s) {
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    if (axis == 0) {
        int nby = (dy + kXLines - 1) / kXLines;
        int64_t blocks = (int64_t)nby * dz;
        sweep_x_kernel<<<(unsigned)blocks, kXLines, 0, s>>>(src, dst, dx, dy, dz, nby, (unsigned)blocks, g.h, dir,
                                                           changed);
    } else {
        int nbx = (dx + 63) / 64;
        int other = (axis == 2) ? dy : dz;
        int64_t blocks = (int64_t)nbx * ((other + kYZRows - 1) / kYZRows);
        dim3 blk(64, kYZRows);
        if (axis == 2)
            sweep_yz_kernel<2><<<(unsigned)blocks, blk, 0, s>>>(src, dst, dx, dy, dz, nbx, (unsigned)blocks, g.h,
                                                               dir, changed);
        else
            sweep_yz_kernel<1><<<(unsigned)blocks, blk, 0, s>>>(src, dst, dx, dy, dz, nbx, (unsigned)blocks, g.h,
                                                               dir, changed);
    }
    return hipGetLastError();
}

}  // namespace

bool lane_sweeps_ok(const Grid& g) {
    const int64_t dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    return dx >= 2 && dy >= 2 && dz >= 2 && dx * dy < (int64_t(1) << 29);
}

hipError_t launch_transpose_xy(const float* in, float* out, const Grid& g, bool to_yx, hipStream_t s) {
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    const int rows = to_yx ? dy : dx, cols = to_yx ? dx : dy;
    const int tr = (rows + 63) / 64, tc = (cols + 63) / 64;
    const int64_t blocks = (int64_t)tr * tc * dz;
    transpose_planes_kernel<<<(unsigned)blocks, 256, 0, s>>>(in, out, rows, cols, tr, tc, (unsigned)blocks);
    return hipGetLastError();
}

hipError_t launch_sweep_yx(const float* src, float* dst, const Grid& g, int dir, int* changed, hipStream_t s) {
    // y-fastest layout: lanes over y (Godunov slot 1), walk x (slot 0, stride dy), cross z (slot 2)
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    return launch_lane<1, 0, 2>(src, dst, dy, dx, dz, dy, (int64_t)dx * dy, g.h, dir, changed, s);
}

hipError_t launch_sweep(const float* src, float* dst, const Grid& g, int axis, int dir, int* changed,
                        hipStream_t s) {
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    if (axis == 0 || !lane_sweeps_ok(g)) return launch_generic_sweep(src, dst, g, axis, dir, changed, s);
    if (axis == 1)  // lanes x (slot 0), walk y (slot 1, stride dx), cross z (slot 2)
        return launch_lane<0, 1, 2>(src, dst, dx, dy, dz, dx, (int64_t)dx * dy, g.h, dir, changed, s);
    // axis 2: lanes x (slot 0), walk z (slot 2, stride dx*dy), cross y (slot 1)
    return launch_lane<0, 2, 1>(src, dst, dx, dz, dy, (int64_t)dx * dy, dx, g.h, dir, changed, s);
}

hipError_t launch_redblack(float* g, const int32_t dims[3], float h, int parity, hipStream_t s) {
    int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    redblack_kernel<<<(unsigned)b, 256, 0, s>>>(g, dims[0], dims[1], dims[2], h, parity);
    return hipGetLastError();
}

}  // namespace rhz
