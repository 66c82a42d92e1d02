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
// The field arrives x fastest.  By default the rounds run on a y-fastest
// copy (one plane transpose before the first round and one after the last):
// the x- and z-sweeps put 64 consecutive y in one wave, so every step is a
// coalesced row access, and the y-sweep walks along the rows themselves, a
// row per lane, in register chunks staged through LDS.  The other layout
// (tuning "fsm_layout" 0) keeps the field x fastest for the y- and z-sweeps
// and transposes around the x-sweeps of every round.  Block order is
// remapped so neighbouring tiles share an XCD's L2.  A generic x-sweep kernel
// covers lattices too large for 32-bit lane offsets.
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
// y-fastest copy.
//
// Addressing costs no vector instructions: every step builds a buffer
// resource (scalar registers) whose base is the wave's current row minus
// A = cs + ws + 1 cells, so the five reads of a step (the two fast-axis
// neighbours, the two cross-line neighbours, the next cell on the line) are
// buffer loads at per-lane byte offsets fixed for the whole walk.  Stores go
// through a resource bounded by the row length: dead lanes (f >= nf) store
// at an offset past the bound, which the hardware drops.  A missing
// neighbour at a grid face reads the opposite neighbour instead:
// fminf(v, v) = fminf(1e30, v) for every stored value (all are <= 1e30), so no
// per-step select is needed (F and C must have at least 2 cells).
//
// The walk runs in groups of STEP steps whose loads are issued together;
// groups never touch a row past the end of the line, and the last < STEP + 1
// steps run one at a time with the end-of-line value 1e30.  ROWS waves (ROWS
// adjacent lines) form a block.  Both are launch-time choices (sweep tuning).
constexpr int kBufWord3 = 0x00020000;  // raw 32-bit buffer access (gfx9 family)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_at(const float* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)bytes, kBufWord3);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, 0, 0));
}
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, 0, 0);
}

// Correctly rounded binary32 square root for 2^-96 <= x < inf in one
// hardware reciprocal square root and four operations: y = rsq(x), s = RN(x y),
// then one Newton step s + RN(x - s^2) (y / 2) with two FMAs.  Compared with
// IEEE sqrt for every binary32 x in that range on the device itself
// (rhz_selftest_sweep_sqrt, sqrt_selftest_kernel below): equal everywhere.
// (The previous form, the hardware sqrt corrected by the residuals of its two
// neighbours, took eight operations.)  For x = 2^-126, used as a stand-in for
// a zero discriminant, it returns 2^-63.
__device__ __forceinline__ float sqrt_rn_normal(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    const float s = x * y;
    const float e = __builtin_fmaf(-s, s, x);
    return __builtin_fmaf(e, 0.5f * y, s);
}

// Wave-uniform choice of square root: the short form when every active lane's
// argument is in its range, the full library form otherwise (zero, tiny).
__device__ __forceinline__ float sqrt_rn_wave(float x) {
    if (__ballot(!(x >= 0x1.0p-96f)) == 0) return sqrt_rn_normal(x);
    return sqrtf(x);
}

// godunov_min with the uniform square roots above; bit-identical results.
__device__ __forceinline__ float godunov_min_w(float cur, float ax, float ay, float az, const SweepConst& K) {
    const float s0 = fminf(ax, fminf(ay, az));
    const float s2 = fmaxf(ax, fmaxf(ay, az));
    if (!(cur > s0 + K.h057) && s2 < K.hbig) return cur;
    const float s1 = ax + ay + az - s0 - s2;
    float u = s0 + K.h;
    if (u > s1) {
        const float d10 = s1 - s0;
        const float disc = K.h2x2 - d10 * d10;
        if (disc >= 0.0f) {
            u = (s0 + s1 + sqrt_rn_wave(disc)) * 0.5f;
            if (u > s2) {
                const float d21 = s2 - s1, d20 = s2 - s0;
                const float disc3 = K.h2x3 - (d10 * d10 + d21 * d21 + d20 * d20) * 0.5f;
                if (disc3 >= 0.0f) u = div3(s0 + s1 + s2 + sqrt_rn_wave(disc3));
            }
        }
    }
    return (u < cur) ? u : cur;
}

template <int SF, int SW, int SC>
__device__ __forceinline__ float lane_step(float cur, float fm, float fp, float cm, float cp, float behind, float next,
                                           const SweepConst& K) {
    float g[3];
    g[SF] = fminf(fm, fp);
    g[SW] = fminf(behind, next);
    g[SC] = fminf(cm, cp);
    // The solve runs only when some lane of the wave can improve (a uniform
    // branch); lanes that skip keep cur, which is what the solve's own exit
    // returns for them.
    const float lo = fminf(g[0], fminf(g[1], g[2]));
    const float hi = fmaxf(g[0], fmaxf(g[1], g[2]));
    const bool need = (cur > lo + K.h057) || !(hi < K.hbig);
    if (__ballot(need) == 0) return cur;
    return godunov_min_w(cur, g[0], g[1], g[2], K);
}

// ---- tile sweep (the default shape) --------------------------------------------
// Every lane walks NF x NC lines at once: NF segments of the fast axis, 64
// cells apart, times NC adjacent cross lines.  Per step a wave needs only its
// own rows: the fast-axis neighbours are the adjacent lanes' cells, moved
// through DPP wave shifts (lane 0 and lane 63 of the outer segments take one
// extra "edge" load), and the cross-line neighbours inside the tile are the
// lane's own registers; only the lines just outside the tile (c0 - 1 and
// c0 + NC) are loaded.  The update is branch-free, so the NF x NC chains
// interleave.  Waves that touch no lattice face skip the face selects.
//
// The square roots use the short correctly rounded form: with
// 2^-30 <= h <= 2^30 every positive discriminant is >= h^2 2^-25 >= 2^-96.
// (disc = RN(a - b) for floats a ~ 2h^2 or 3h^2 and b >= 0: either b <= a/2
// and disc >= a/2, or a - b is exact and a multiple of b's unit in the last
// place, > b 2^-24 > a 2^-25.)  A negative discriminant is never selected.
// A zero one is raised to 2^-126 (one max instead of a compare and a select):
// its root, about 2^-63, is then added to a sum >= h sqrt(2) >= 2^-30.5 (zero
// 2-D discriminant: s1 - s0 = sqrt(2) h; zero 3-D one: s2 - s0 >= sqrt(2) h),
// less than half that sum's unit in the last place, so the sum and the
// candidate are those of root 0 bit for bit.
__device__ __forceinline__ float sqrt_disc(float disc) { return sqrt_rn_normal(fmaxf(disc, 0x1.0p-126f)); }

// x / 3 correctly rounded in two operations, RN(x zh + RN(x zl)) with
// zh = RN(1/3) and zl = RN(1/3 - zh) (multiplication by a reciprocal split in
// two, with one FMA; Brisebarre and Muller, IEEE TC 2008).  Compared with IEEE
// division over every non-negative finite binary32 value: equal for every
// x > 2^-123 (below that x zl underflows).  The sweeps' 3-D numerator is
// s0 + s1 + s2 + root >= h / 2 >= 2^-31 wherever it is selected.
__device__ __forceinline__ float div3_split(float x) {
    return __builtin_fmaf(x, 0x1.555556p-2f, x * -0x1.555556p-27f);
}

// godunov_min with every candidate root computed and then selected: the roots
// the branchy form would not reach are discarded by the selects, and where its
// early exit returns cur the full formula returns cur too, so the results are
// identical bit for bit.  godunov_cand is the candidate before the final
// min with cur: callers that track changes take u < cur once, for the select
// and for a wave mask (a scalar OR per step instead of two vector ops).
// The 3-D discriminant h2x3 - RN(q 0.5) is one FMA, RN(h2x3 - q 0.5): q 0.5
// is exact for q >= 2^-125, and below that both forms round to h2x3.
__device__ __forceinline__ float godunov_cand(float a0, float a1, float a2, const SweepConst& K) {
    const float s0 = fminf(a0, fminf(a1, a2));
    const float s2 = fmaxf(a0, fmaxf(a1, a2));
    const float s1 = a0 + a1 + a2 - s0 - s2;
    const float u1 = s0 + K.h;
    const float d10 = s1 - s0;
    const float disc = K.h2x2 - d10 * d10;
    const float u2 = (s0 + s1 + sqrt_disc(disc)) * 0.5f;
    const float d21 = s2 - s1, d20 = s2 - s0;
    const float disc3 = __builtin_fmaf(d10 * d10 + d21 * d21 + d20 * d20, -0.5f, K.h2x3);
    const float u3 = div3_split(s0 + s1 + s2 + sqrt_disc(disc3));
    const bool two = (u1 > s1) && (disc >= 0.0f);
    const bool three = two && (u2 > s2) && (disc3 >= 0.0f);
    return three ? u3 : (two ? u2 : u1);
}
__device__ __forceinline__ float godunov_select(float cur, float a0, float a1, float a2, const SweepConst& K) {
    const float u = godunov_cand(a0, a1, a2, K);
    return (u < cur) ? u : cur;
}

// Wave shifts by one lane (DPP): lane i takes lane i - 1 (shr) or i + 1 (shl);
// the lane with no source keeps `old`.  ror / rol wrap around the wave.
__device__ __forceinline__ float wave_shr1(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float wave_shl1(float old, float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float wave_ror1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x13c, 0xf, 0xf, false));
}
__device__ __forceinline__ float wave_rol1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x134, 0xf, 0xf, false));
}

template <int SF, int SW, int SC, int DIR, int STEP, int NF, int NC, bool FACE>
__device__ __forceinline__ bool tile_walk(const float* __restrict__ src, float* __restrict__ dst, int nf, int nw,
                                          int nc, int64_t ws, int64_t cs, int bf, int c0, const SweepConst& K) {
    const int lane = threadIdx.x;
    const int64_t A = cs + ws + 1;  // resource base = row - A: every offset below is >= 0
    const bool need_lo = c0 > 0;
    const bool need_hi = c0 + NC <= nc - 1;
    bool lo_face[NC], hi_face[NC];
    int64_t coff[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const int c = (c0 + i < nc) ? c0 + i : nc - 1;  // lines past the end repeat the last one, unstored
        lo_face[i] = c == 0;
        hi_face[i] = c == nc - 1;
        coff[i] = (int64_t)(c - c0) * cs;
    }
    // byte offsets of the NEXT row's cells (the loads run one step ahead)
    uint32_t v_nx[NC][NF], v_st[NC][NF], v_lo[NF], v_hi[NF], v_ed[NC];
    bool hasl[NF], hasr[NF];
#pragma unroll
    for (int j = 0; j < NF; ++j) {
        const int f = (bf * NF + j) * 64 + lane;
        const bool live = f < nf;
        const int fl = live ? f : nf - 1;
        hasl[j] = fl > 0;
        hasr[j] = fl < nf - 1;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            v_nx[i][j] = (uint32_t)((fl + coff[i] + (int64_t)DIR * ws + A) * 4);
            v_st[i][j] = (live && c0 + i < nc) ? (uint32_t)((coff[i] + f) * 4) : 0x7ffffff0u;
        }
        v_lo[j] = (uint32_t)((fl - cs + (int64_t)DIR * ws + A) * 4);
        v_hi[j] = (uint32_t)((fl + (int64_t)NC * cs + (int64_t)DIR * ws + A) * 4);
    }
    {
        // lane 63 reads the cell right of the last segment, every other lane the
        // cell left of the first (clamped into the row; unused at a face)
        const int f0 = bf * NF * 64;
        const int fr = f0 + NF * 64 - 1;
        const int fe = (lane == 63) ? ((fr < nf - 1) ? fr + 1 : nf - 1) : ((f0 > 0) ? f0 - 1 : 0);
#pragma unroll
        for (int i = 0; i < NC; ++i) v_ed[i] = (uint32_t)((fe + coff[i] + (int64_t)DIR * ws + A) * 4);
    }
    const uint32_t nrec = (uint32_t)((2 * A + (int64_t)NC * cs + nf + 64 * NF) * 4);
    const uint32_t nrows = (uint32_t)(((int64_t)(NC - 1) * cs + nf) * 4);
    const int64_t step = (int64_t)DIR * ws;
    int64_t row = (int64_t)c0 * cs + (int64_t)((DIR > 0) ? 0 : nw - 1) * ws;
    float behind[NC][NF], cur[NC][NF], edge[NC], hlo[NF], hhi[NF];
    unsigned long long chgm = 0;  // lanes that lowered a cell
    {
        const __amdgpu_buffer_rsrc_t r0 = rsrc_at(src + row - A, nrec);
        const uint32_t back = (uint32_t)(step * 4);
#pragma unroll
        for (int i = 0; i < NC; ++i) {
#pragma unroll
            for (int j = 0; j < NF; ++j) {
                behind[i][j] = kFar;
                cur[i][j] = bload(r0, v_nx[i][j] - back);
            }
            edge[i] = bload(r0, v_ed[i] - back);
        }
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            hlo[j] = need_lo ? bload(r0, v_lo[j] - back) : 0.0f;
            hhi[j] = need_hi ? bload(r0, v_hi[j] - back) : 0.0f;
        }
    }
    // one step: update every cell of the row from cur / edge / halo and the
    // next row nx, store, and shift the next row in
    auto advance = [&](const float (&nx)[NC][NF], const float (&ed)[NC], const float (&lo)[NF],
                       const float (&hi)[NF]) {
        const __amdgpu_buffer_rsrc_t w = rsrc_at(dst + row, nrows);
        float nv[NC][NF];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
#pragma unroll
            for (int j = 0; j < NF; ++j) {
                const float left = (j == 0) ? wave_shr1(edge[i], cur[i][0])
                                            : wave_ror1((lane == 63) ? cur[i][j > 0 ? j - 1 : 0] : cur[i][j]);
                const float right = (j == NF - 1)
                                        ? wave_shl1(edge[i], cur[i][NF - 1])
                                        : wave_rol1((lane == 0) ? cur[i][j < NF - 1 ? j + 1 : 0] : cur[i][j]);
                float fm = left, fp = right;
                float cm = (i > 0) ? cur[i > 0 ? i - 1 : 0][j] : hlo[j];
                float cp = (i < NC - 1) ? cur[i < NC - 1 ? i + 1 : 0][j] : hhi[j];
                if (FACE) {  // a missing neighbour is replaced by the opposite one
                    fm = hasl[j] ? left : right;
                    fp = hasr[j] ? right : left;
                    if (lo_face[i]) cm = cp;
                    if (hi_face[i]) cp = cm;
                }
                float g[3];
                g[SF] = fminf(fm, fp);
                g[SW] = fminf(behind[i][j], nx[i][j]);
                g[SC] = fminf(cm, cp);
                const float u = godunov_cand(g[0], g[1], g[2], K);
                const bool lower = u < cur[i][j];
                nv[i][j] = lower ? u : cur[i][j];
                chgm |= __ballot(lower);
            }
        }
#pragma unroll
        for (int i = 0; i < NC; ++i) {
#pragma unroll
            for (int j = 0; j < NF; ++j) {
                bstore(w, v_st[i][j], nv[i][j]);
                behind[i][j] = nv[i][j];
                cur[i][j] = nx[i][j];
            }
            edge[i] = ed[i];
        }
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            hlo[j] = lo[j];
            hhi[j] = hi[j];
        }
        row += step;
    };
    int s0 = 0;
    // full groups: every step of the group has a next row on the line
    for (; s0 + STEP < nw; s0 += STEP) {
        float nx[STEP][NC][NF], ed[STEP][NC], lo[STEP][NF], hi[STEP][NF];
#pragma unroll
        for (int k = 0; k < STEP; ++k) {
            const __amdgpu_buffer_rsrc_t r = rsrc_at(src + (row + k * step) - A, nrec);
#pragma unroll
            for (int i = 0; i < NC; ++i) {
#pragma unroll
                for (int j = 0; j < NF; ++j) nx[k][i][j] = bload(r, v_nx[i][j]);
                ed[k][i] = bload(r, v_ed[i]);
            }
#pragma unroll
            for (int j = 0; j < NF; ++j) {
                lo[k][j] = need_lo ? bload(r, v_lo[j]) : 0.0f;
                hi[k][j] = need_hi ? bload(r, v_hi[j]) : 0.0f;
            }
        }
#pragma unroll
        for (int k = 0; k < STEP; ++k) advance(nx[k], ed[k], lo[k], hi[k]);
    }
    // the last steps one at a time; past the end of the line next = 1e30
    for (; s0 < nw; ++s0) {
        const bool more = s0 + 1 < nw;
        const __amdgpu_buffer_rsrc_t r = rsrc_at(src + row - A, nrec);
        float nx[NC][NF], ed[NC], lo[NF], hi[NF];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
#pragma unroll
            for (int j = 0; j < NF; ++j) nx[i][j] = more ? bload(r, v_nx[i][j]) : kFar;
            ed[i] = more ? bload(r, v_ed[i]) : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            lo[j] = (more && need_lo) ? bload(r, v_lo[j]) : 0.0f;
            hi[j] = (more && need_hi) ? bload(r, v_hi[j]) : 0.0f;
        }
        advance(nx, ed, lo, hi);
    }
    return chgm != 0;
}

// Block = 4 waves; wave (bf, g) takes fast-axis cells [64 NF bf, 64 NF (bf + 1))
// of cross lines [NC g, NC g + NC).
template <int SF, int SW, int SC, int DIR, int STEP, int NF, int NC>
__global__ __launch_bounds__(256) void sweep_tile_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                         int nf, int nw, int nc, int64_t ws, int64_t cs, int nbf,
                                                         unsigned nblocks, float h, int* __restrict__ changed) {
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const int bf = (int)(lb % (unsigned)nbf);
    const int c0 = __builtin_amdgcn_readfirstlane(((int)(lb / (unsigned)nbf) * 4 + (int)threadIdx.y) * NC);
    if (c0 >= nc) return;
    const SweepConst K = sweep_const(h);
    const bool face = c0 == 0 || c0 + NC >= nc || bf == 0 || (bf + 1) * NF * 64 >= nf;
    const bool chg = face ? tile_walk<SF, SW, SC, DIR, STEP, NF, NC, true>(src, dst, nf, nw, nc, ws, cs, bf, c0, K)
                          : tile_walk<SF, SW, SC, DIR, STEP, NF, NC, false>(src, dst, nf, nw, nc, ws, cs, bf, c0, K);
    if (__ballot(chg) != 0 && threadIdx.x == 0) atomicOr(changed, 1);
}

// ---- y-sweep on the y-fastest copy (the walk runs along the contiguous axis) ----
// Layout y-fastest: cell (x, y, z) at z * plane + x * pitch + y, rows padded to
// a multiple of 32 floats (128 B).  One wave = 64 consecutive x (one y-row per
// lane) of one z; a block = 4 consecutive z.  Each lane walks its row in
// chunks of CH cells held in registers: its own row, the rows at z - 1 and
// z + 1 (the opposite one at a face: fminf(v, v) = fminf(1e30, v) for every
// stored value), and for lanes 0 and 63 the row just outside the wave in x.
// The x neighbours of a cell are the adjacent lanes' cells of the same step
// (DPP wave shifts), the y neighbours the new value behind the walker and the
// old value ahead (the first cell of the next chunk read ahead), so the
// update reads exactly what the pinned schedule prescribes: results equal
// the y-sweep on the x-fastest layout bit for bit.  Chunk indices are static:
// the one ragged chunk of a row (at its far end) guards each step.
// One cell of the walk: x neighbours from the adjacent lanes (DPP; lane 0 and
// lane 63 keep their edge-row value e), y neighbours behind / next, z
// neighbours m / p.  Branch-free (x, y, z slots in order): one chain per lane
// leaves no room for a wave-uniform branch every step.
__device__ __forceinline__ float run_cell(float cur, float mp, float e, float behind, float next, bool xlo_face,
                                          bool xhi_face, const SweepConst& K) {
    const float l = wave_shr1(e, cur);
    const float r = wave_shl1(e, cur);
    const float fm = xlo_face ? r : l;
    const float fp = xhi_face ? l : r;
    return godunov_select(cur, fminf(fm, fp), fminf(behind, next), mp, K);
}

template <int DIR, int CH>
__device__ __forceinline__ void run_chunk(float (&o)[CH], const float (&m)[CH], const float (&p)[CH],
                                          const float (&e)[CH], float look, int nvalid, bool xlo_face,
                                          bool xhi_face, float& behind, bool& chg, const SweepConst& K) {
#pragma unroll
    for (int t = 0; t < CH; ++t) {
        const int k = DIR > 0 ? t : CH - 1 - t;
        if (k < nvalid) {  // wave-uniform
            const float cur = o[k];
            float next;
            if (DIR > 0)
                next = (k + 1 < nvalid) ? o[k + 1 < CH ? k + 1 : k] : look;
            else
                next = (k > 0) ? o[k > 0 ? k - 1 : 0] : look;
            const float nv = run_cell(cur, fminf(m[k], p[k]), e[k], behind, next, xlo_face, xhi_face, K);
            chg |= nv < cur;
            o[k] = nv;
            behind = nv;
        }
    }
}

// A wave's rows move between global memory and registers through its own LDS
// tile (64 rows x CH cells, one padding float per row: a lane's row and a
// column of 8-row groups are both conflict-free).  Global accesses are then
// coalesced: instruction i of a wave covers rows 8 i .. 8 i + 7, CH / 8 cells
// of each per lane group.  The wave-scope fences keep the compiler from
// moving a lane's LDS read of another lane's write across them.
template <int CH>
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// rows x0 + r (r = 0..63, clamped to the last row) of `base` (row r at
// base + r * pitch), cells [y0, y0 + CH), into v[] of lane r
template <int CH>
__device__ __forceinline__ void stage_in(float (&v)[CH], float* __restrict__ tile, const float* __restrict__ base,
                                         int64_t pitch, int rmax, int y0) {
    constexpr int Q = CH / 4;      // float4 per row
    constexpr int RPI = 64 / Q;    // rows per instruction
    const int lane = threadIdx.x;
    const int rr = lane / Q, qq = lane % Q;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int r = i * RPI + rr;
        const float4 a = reinterpret_cast<const float4*>(base + (int64_t)min(r, rmax) * pitch + y0)[qq];
        float* t = tile + r * (CH + 1) + 4 * qq;
        t[0] = a.x;
        t[1] = a.y;
        t[2] = a.z;
        t[3] = a.w;
    }
    wave_lds_sync<CH>();
#pragma unroll
    for (int k = 0; k < CH; ++k) v[k] = tile[lane * (CH + 1) + k];
    wave_lds_sync<CH>();
}

// v[] of lane r to rows x0 + r (r < nlive) of `base`, cells [y0, y0 + CH)
template <int CH>
__device__ __forceinline__ void stage_out(const float (&v)[CH], float* __restrict__ tile, float* __restrict__ base,
                                          int64_t pitch, int nlive, int y0) {
    constexpr int Q = CH / 4;
    constexpr int RPI = 64 / Q;
    const int lane = threadIdx.x;
    const int rr = lane / Q, qq = lane % Q;
#pragma unroll
    for (int k = 0; k < CH; ++k) tile[lane * (CH + 1) + k] = v[k];
    wave_lds_sync<CH>();
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int r = i * RPI + rr;
        const float* t = tile + r * (CH + 1) + 4 * qq;
        if (r < nlive) reinterpret_cast<float4*>(base + (int64_t)r * pitch + y0)[qq] = make_float4(t[0], t[1], t[2], t[3]);
    }
    wave_lds_sync<CH>();
}

template <int DIR, int CH>
__global__ __launch_bounds__(256) void sweep_run_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                        int nx, int ny, int nz, int64_t pitch, int64_t plane,
                                                        int nbx, unsigned nblocks, float h,
                                                        int* __restrict__ changed) {
    __shared__ float tiles[4][64 * (CH + 1)];
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const int bx = (int)(lb % (unsigned)nbx);
    const int z = __builtin_amdgcn_readfirstlane((int)(lb / (unsigned)nbx) * 4 + (int)threadIdx.y);
    if (z >= nz) return;
    float* tile = tiles[threadIdx.y];
    const int lane = threadIdx.x;
    const int x0 = bx * 64;
    const int x = x0 + lane;
    const int nlive = min(64, nx - x0);  // rows of this wave inside the lattice
    const int rmax = nlive - 1;
    const int xl = x < nx ? x : nx - 1;
    int xe = xl;
    if (lane == 0 && xl > 0) xe = xl - 1;
    if (lane == 63 && xl < nx - 1) xe = xl + 1;
    const bool xlo_face = x == 0, xhi_face = x >= nx - 1;
    const int64_t zo = (int64_t)z * plane;
    const float* own = src + zo + (int64_t)x0 * pitch;  // row r of the wave at own + r * pitch
    const float* zm = own + (z > 0 ? -plane : plane);
    const float* zp = own + (z < nz - 1 ? plane : -plane);
    const float* ed = src + zo + (int64_t)xe * pitch;
    const float* mine = src + zo + (int64_t)xl * pitch;
    float* out = dst + zo + (int64_t)x0 * pitch;
    const bool edge_lane = lane == 0 || lane == 63;
    const SweepConst K = sweep_const(h);
    float behind = kFar;
    bool chg = false;
    const int nch = (ny + CH - 1) / CH;
    for (int c = 0; c < nch; ++c) {
        const int ci = DIR > 0 ? c : nch - 1 - c;
        const int y0 = ci * CH;
        const int nvalid = min(CH, ny - y0);
        float o[CH], m[CH], p[CH], e[CH];
        stage_in<CH>(o, tile, own, pitch, rmax, y0);
        stage_in<CH>(m, tile, zm, pitch, rmax, y0);
        stage_in<CH>(p, tile, zp, pitch, rmax, y0);
        if (edge_lane) {
#pragma unroll
            for (int q = 0; q < CH / 4; ++q) {
                const float4 a = reinterpret_cast<const float4*>(ed + y0)[q];
                e[4 * q] = a.x; e[4 * q + 1] = a.y; e[4 * q + 2] = a.z; e[4 * q + 3] = a.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < CH; ++q) e[q] = 0.0f;
        }
        // the first cell past the chunk along the walk (1e30 past the row's end)
        const int yl = DIR > 0 ? y0 + CH : y0 - 1;
        const float look = (yl >= 0 && yl < ny) ? mine[yl] : kFar;
        run_chunk<DIR, CH>(o, m, p, e, look, nvalid, xlo_face, xhi_face, behind, chg, K);
        stage_out<CH>(o, tile, out, pitch, nlive, y0);
    }
    if (__ballot(chg) != 0 && threadIdx.x == 0) atomicOr(changed, 1);
}

// ---- y-sweep on the y-fastest copy, prefetching ----------------------------------
// Same walk and arithmetic as sweep_run_kernel (so the same results bit for
// bit), organised so that each plane's rows are fetched once per block and
// no load is waited for inside a chunk:
//  * every lane walks C rows at once, one in each of C adjacent planes;
//  * a block is W waves over the planes z0 .. z0 + W C - 1 and 64 rows: the
//    wave's 62 rows x0 .. x0 + 61 on lanes 1 .. 62, and on lanes 0 and 63 the
//    rows just outside them, walked only to hand their old values to lanes 1
//    and 62 (DPP) and never stored;
//  * per chunk of CH cells, the rows of the W C + 2 planes z0 - 1 .. z0 + W C
//    are staged once in LDS and shared (a row's z - 1 and z + 1 neighbours are
//    the rows of the adjacent planes): sweep_run_kernel's waves each fetched
//    their own z - 1 and z + 1 rows, 2.2x the field's bytes per sweep past L2;
//  * the chunk after next is loaded into registers (coalesced float4, four
//    lanes per 64-byte row piece) while a chunk is computed, and written into
//    LDS one chunk later; a chunk's last cell along the walk, whose `next` is
//    the first cell of the following chunk, is finished once that chunk is in
//    registers, so no wait falls inside a chunk;
//  * results go to memory through a per-plane LDS tile, so that a store
//    instruction writes whole 16-byte-aligned row pieces (a lane storing its
//    own row's pieces straight from registers measured 40% slower);
//  * LDS rows are padded to CH + 4 floats, so 16-byte reads of a lane's own
//    row touch distinct banks, and so do the row pieces of the staging and
//    the stores, whose lanes pair rows 4 apart (with adjacent rows the
//    counters showed 2.4 times the bank-conflict cycles; the sweep's time
//    moved by 0.5%: LDS is not what bounds it).
// Measured on the 883 x 823 x 923 lattice and not adopted: direct global->LDS
// loads (global_load_lds_dwordx4; a lane's 16 bytes became a request of its
// own), two planes per wave (C = 2: about 200 VGPRs, two waves per SIMD), and
// whole 128-byte lines per row piece (32-cell chunks, stores straight from
// registers: 3.2-3.7 GB read per sweep instead of 7.0, but 214-256 VGPRs and a
// slower sweep).  The kernel is only marginally faster than sweep_run_kernel:
// both are bound by the single dependency chain per lane at the occupancy
// that LDS and registers allow.
// Rows and planes past the lattice edge are loaded clamped into the lattice
// and never stored; only stored cells count as changed.
// The minimum of a cell's two x neighbours in the prefetching kernel, whose
// lanes 0 and 63 are never stored: their DPP sources may be anything.  FACE
// (the wave holds row 0 or row nx - 1): a missing neighbour is replaced by the
// opposite one.
template <bool FACE>
__device__ __forceinline__ float run_xmin(float cur, bool xlo_face, bool xhi_face) {
    if (FACE) {
        const float l = wave_shr1(cur, cur);
        const float r = wave_shl1(cur, cur);
        return fminf(xlo_face ? r : l, xhi_face ? l : r);
    }
    const float l = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(cur), 0x138, 0xf, 0xf, true));
    const float r = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(cur), 0x130, 0xf, 0xf, true));
    return fminf(l, r);
}

// One cell of the prefetching kernel's walk; lanes that lower it go into chgm.
template <bool FACE>
__device__ __forceinline__ float run_cellm(float cur, float mp, float behind, float next, bool xlo_face,
                                           bool xhi_face, unsigned long long& chgm, const SweepConst& K) {
    const float u = godunov_cand(run_xmin<FACE>(cur, xlo_face, xhi_face), fminf(behind, next), mp, K);
    const bool lower = u < cur;
    chgm |= __ballot(lower);
    return lower ? u : cur;
}

// C chains per lane: a chunk of C rows at once, steps interleaved.  FULL: a
// whole chunk (no guard per step); SKIPLAST: the chunk's last cell along the
// walk is left for the caller (its `next` is in the following chunk).
template <int DIR, int CH, int C, bool FULL, bool SKIPLAST, bool FACE>
__device__ __forceinline__ void run_chunks(float (&o)[C][CH], const float (&mp)[C][CH], int nvalid_, bool xlo_face,
                                           bool xhi_face, float (&behind)[C], unsigned long long (&chgm)[C],
                                           const SweepConst& K) {
    const int nvalid = FULL ? CH : nvalid_;
#pragma unroll
    for (int t = 0; t < (SKIPLAST ? CH - 1 : CH); ++t) {
        const int k = DIR > 0 ? t : CH - 1 - t;
        if (k < nvalid) {  // wave-uniform
#pragma unroll
            for (int c = 0; c < C; ++c) {
                float next;
                if (DIR > 0)
                    next = (k + 1 < nvalid) ? o[c][k + 1 < CH ? k + 1 : k] : kFar;
                else
                    next = (k > 0) ? o[c][k > 0 ? k - 1 : 0] : kFar;
                const float nv = run_cellm<FACE>(o[c][k], mp[c][k], behind[c], next, xlo_face, xhi_face, chgm[c], K);
                o[c][k] = nv;
                behind[c] = nv;
            }
        }
    }
}

// A chunk of the prefetching kernel: the walk to the row's end (!more) or to
// the chunk's last cell; only a y-descending sweep starts on the ragged chunk,
// and its last cell (k = 0) is whole.
template <int DIR, int CH, int C, bool FACE>
__device__ __forceinline__ void run_chunk_any(float (&o)[C][CH], const float (&mp)[C][CH], int nvalid, bool more,
                                              bool xlo_face, bool xhi_face, float (&behind)[C],
                                              unsigned long long (&chgm)[C], const SweepConst& K) {
    if (!more) {
        if (nvalid == CH) run_chunks<DIR, CH, C, true, false, FACE>(o, mp, CH, xlo_face, xhi_face, behind, chgm, K);
        else run_chunks<DIR, CH, C, false, false, FACE>(o, mp, nvalid, xlo_face, xhi_face, behind, chgm, K);
    } else {
        if (nvalid == CH) run_chunks<DIR, CH, C, true, true, FACE>(o, mp, CH, xlo_face, xhi_face, behind, chgm, K);
        else run_chunks<DIR, CH, C, false, true, FACE>(o, mp, nvalid, xlo_face, xhi_face, behind, chgm, K);
    }
}

template <int DIR, int CH, int W, int C, bool ZMAJ = false>
__global__ __launch_bounds__(64 * W) void sweep_runp_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                            int nx, int ny, int nz, int64_t pitch, int64_t plane,
                                                            int nbx, unsigned nblocks, float h,
                                                            int* __restrict__ changed) {
    constexpr int Q = CH / 4;     // float4 per row piece
    constexpr int RPI = 64 / Q;   // rows per load instruction
    constexpr int RS = CH + 4;    // LDS row stride (floats)
    constexpr int P = W * C;      // planes per block
    constexpr int NS = P + 2;     // staged planes z0 - 1 .. z0 + P
    constexpr int NL = NS * Q;    // load instructions per chunk (Q per plane)
    constexpr int NLW = NL / W;   // ... per wave
    static_assert(NL % W == 0, "the waves share the loads evenly");
    constexpr int PK = DIR > 0 ? CH - 1 : 0;  // a chunk's last cell along the walk
    __shared__ __align__(16) float in_t[NS][64 * RS];
    __shared__ __align__(16) float out_t[P][64 * RS];
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const unsigned nzg = nblocks / (unsigned)nbx;  // plane groups
    // ZMAJ: consecutive logical blocks (one XCD) are adjacent plane groups,
    // which stage each other's boundary planes
    const int bx = ZMAJ ? (int)(lb / nzg) : (int)(lb % (unsigned)nbx);
    const int z0 = (ZMAJ ? (int)(lb % nzg) : (int)(lb / (unsigned)nbx)) * P;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int lane = (int)threadIdx.x;
    const int x0 = bx * 62;
    const int x = x0 - 1 + lane;  // lanes 0 and 63: the rows beside the wave's 62
    const bool xlo_face = x == 0, xhi_face = x == nx - 1;
    const bool mine = lane >= 1 && lane <= 62 && x < nx;
    const bool face = x0 == 0 || x0 + 62 >= nx;  // a stored lane holds row 0 or row nx - 1
    bool plive[C];
    int os[C], ms[C], ps[C];
    float* out[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int z = z0 + w * C + c;
        const int zc = z < nz ? z : nz - 1;
        plive[c] = z < nz;
        os[c] = w * C + c + 1;                         // this plane's staged rows
        ms[c] = zc > 0 ? os[c] - 1 : os[c] + 1;        // z - 1 (the opposite plane at a face)
        ps[c] = zc < nz - 1 ? os[c] + 1 : os[c] - 1;   // z + 1
        out[c] = dst + (int64_t)zc * plane;
    }
    // load instruction t of this wave: flattened j = w + t W = s Q + i (plane s,
    // instruction i); lane -> row i RPI + lr, quad lane % Q.  With 16-cell
    // chunks the 8 lanes that LDS serves together (8 g .. 8 g + 7) take two
    // rows 4 apart: at a stride of 20 floats their 16-byte pieces then fill 8
    // distinct 4-bank groups (two adjacent rows would share one)
    const int lr = Q == 4 ? ((lane >> 2) & 1) * 4 + ((lane >> 3) & 3) + ((lane >> 5) << 3) : lane / Q;
    const int lq = lane % Q;
    const float* lsrc[NLW];
    int lofs[NLW];
#pragma unroll
    for (int t = 0; t < NLW; ++t) {
        const int j = w + t * W;
        const int sp = j / Q, i = j % Q;
        const int zs = min(max(z0 - 1 + sp, 0), nz - 1);
        const int r = i * RPI + lr;
        const int xr = min(max(x0 - 1 + r, 0), nx - 1);
        lsrc[t] = src + (int64_t)zs * plane + (int64_t)xr * pitch + 4 * lq;
        lofs[t] = sp * 64 * RS + r * RS + 4 * lq;
    }
    typedef float vf4 __attribute__((ext_vector_type(4)));  // a plain vector: stays in registers
    vf4 pf[NLW];
#define RHZ_RUNP_LOAD(Y0)                                                                   \
    _Pragma("unroll") for (int t = 0; t < NLW; ++t) pf[t] = *reinterpret_cast<const vf4*>(lsrc[t] + (Y0));
#define RHZ_RUNP_STAGE()                                                                    \
    _Pragma("unroll") for (int t = 0; t < NLW; ++t) *reinterpret_cast<vf4*>(&in_t[0][0] + lofs[t]) = pf[t];
    // out_t[plane] -> rows x0 .. x0 + 61 (lanes 1 .. 62) that lie in the lattice
    auto store = [&](int y0) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            if (!plive[c]) continue;  // wave-uniform
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                const int r = i * RPI + lr;
                const int xr = x0 - 1 + r;
                if (r >= 1 && r <= 62 && xr < nx)
                    *reinterpret_cast<float4*>(out[c] + (int64_t)xr * pitch + y0 + 4 * lq) =
                        *reinterpret_cast<const float4*>(&out_t[w * C + c][r * RS + 4 * lq]);
            }
        }
    };
    const SweepConst K = sweep_const(h);
    const int nch = (ny + CH - 1) / CH;
    auto chunk_y0 = [&](int ck) { return (DIR > 0 ? ck : nch - 1 - ck) * CH; };
    RHZ_RUNP_LOAD(chunk_y0(0))
    RHZ_RUNP_STAGE()
    if (nch > 1) {
        RHZ_RUNP_LOAD(chunk_y0(1))
    }
    float behind[C];
    unsigned long long chgm[C];  // lanes that lowered a cell, per plane
    float pc[C], pmp[C];  // the previous chunk's last cell: old value, z-neighbour minimum
#pragma unroll
    for (int c = 0; c < C; ++c) {
        behind[c] = kFar;
        chgm[c] = 0;
        pc[c] = pmp[c] = 0.0f;
    }
    for (int ck = 0; ck < nch; ++ck) {
        const int y0 = chunk_y0(ck);
        const bool more = ck + 1 < nch;
        __syncthreads();  // every wave's rows of this chunk are staged
        float o[C][CH], mp[C][CH];
#pragma unroll
        for (int c = 0; c < C; ++c) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const float4 a = *reinterpret_cast<const float4*>(&in_t[os[c]][lane * RS + 4 * q]);
                const float4 b = *reinterpret_cast<const float4*>(&in_t[ms[c]][lane * RS + 4 * q]);
                const float4 d = *reinterpret_cast<const float4*>(&in_t[ps[c]][lane * RS + 4 * q]);
                o[c][4 * q] = a.x; o[c][4 * q + 1] = a.y; o[c][4 * q + 2] = a.z; o[c][4 * q + 3] = a.w;
                mp[c][4 * q] = fminf(b.x, d.x); mp[c][4 * q + 1] = fminf(b.y, d.y);
                mp[c][4 * q + 2] = fminf(b.z, d.z); mp[c][4 * q + 3] = fminf(b.w, d.w);
            }
        }
        __syncthreads();  // the tiles are free for the next chunk
        if (more) {  // the next chunk (loaded a chunk ago) into LDS, the one after into registers
            RHZ_RUNP_STAGE()
        }
        if (ck + 2 < nch) {
            RHZ_RUNP_LOAD(chunk_y0(ck + 2))
        }
        // keep the stores below: the staging waits for its loads, and a store
        // issued before it would be waited for as well
        __builtin_amdgcn_sched_barrier(0);
        if (ck > 0) {  // finish the previous chunk: its last cells, then the chunk goes out
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const float next = DIR > 0 ? o[c][0] : o[c][CH - 1];
                const float nv = face ? run_cellm<true>(pc[c], pmp[c], behind[c], next, xlo_face, xhi_face, chgm[c], K)
                                      : run_cellm<false>(pc[c], pmp[c], behind[c], next, xlo_face, xhi_face, chgm[c], K);
                behind[c] = nv;
                out_t[w * C + c][lane * RS + PK] = nv;
            }
            wave_lds_sync<CH>();
            store(chunk_y0(ck - 1));
        }
        const int nvalid = min(CH, ny - y0);
        if (face) run_chunk_any<DIR, CH, C, true>(o, mp, nvalid, more, xlo_face, xhi_face, behind, chgm, K);
        else run_chunk_any<DIR, CH, C, false>(o, mp, nvalid, more, xlo_face, xhi_face, behind, chgm, K);
        wave_lds_sync<CH>();  // this wave's stores have read out_t
#pragma unroll
        for (int c = 0; c < C; ++c) {
#pragma unroll
            for (int q = 0; q < Q; ++q)
                *reinterpret_cast<float4*>(&out_t[w * C + c][lane * RS + 4 * q]) =
                    make_float4(o[c][4 * q], o[c][4 * q + 1], o[c][4 * q + 2], o[c][4 * q + 3]);
            pc[c] = o[c][PK];
            pmp[c] = mp[c][PK];
        }
        wave_lds_sync<CH>();
    }
    store(chunk_y0(nch - 1));
#undef RHZ_RUNP_LOAD
#undef RHZ_RUNP_STAGE
    unsigned long long any = 0;
#pragma unroll
    for (int c = 0; c < C; ++c)
        if (plive[c]) any |= chgm[c];
    if ((any & __ballot(mine)) != 0 && lane == 0) atomicOr(changed, 1);
}



// The job (lattice) that logical block lb belongs to: starts[j] <= lb <
// starts[j + 1] (starts[0] = 0, starts[njobs] = the grid size).
__device__ __forceinline__ int batch_job(const unsigned* __restrict__ starts, int njobs, unsigned lb) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= lb) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One directional sweep of many lattices in one launch: the blocks of job j
// are the blocks sweep_tile_kernel would launch for it alone, so every
// lattice gets exactly its own sweep.  Round r of job j runs only if round
// r - 1 lowered a voxel of that lattice (flags[r - 1]); a round after one
// without change would leave the field as it is.  Changes go to flags[r].
template <int SF, int SW, int SC, int DIR, int STEP, int NF, int NC>
__global__ __launch_bounds__(256) void sweep_tile_batch_kernel(const SweepJob* __restrict__ jobs,
                                                               const unsigned* __restrict__ starts, int njobs,
                                                               unsigned nblocks, int r) {
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const int j = __builtin_amdgcn_readfirstlane(batch_job(starts, njobs, lb));
    const SweepJob J = jobs[j];
    if (r > 0 && J.flags[r - 1] == 0) return;
    const unsigned local = lb - starts[j];
    const int bf = (int)(local % (unsigned)J.nbf);
    const int c0 = __builtin_amdgcn_readfirstlane(((int)(local / (unsigned)J.nbf) * 4 + (int)threadIdx.y) * NC);
    if (c0 >= J.nc) return;
    const SweepConst K = sweep_const(J.h);
    const bool face = c0 == 0 || c0 + NC >= J.nc || bf == 0 || (bf + 1) * NF * 64 >= J.nf;
    const bool chg =
        face ? tile_walk<SF, SW, SC, DIR, STEP, NF, NC, true>(J.src, J.dst, J.nf, J.nw, J.nc, J.ws, J.cs, bf, c0, K)
             : tile_walk<SF, SW, SC, DIR, STEP, NF, NC, false>(J.src, J.dst, J.nf, J.nw, J.nc, J.ws, J.cs, bf, c0, K);
    if (__ballot(chg) != 0 && threadIdx.x == 0) atomicOr(J.flags + r, 1);
}

// One wave walks line row c, cells bf * 64 + lane of the fast axis (the
// one-line walker).  Returns whether this lane lowered a cell.
template <int SF, int SW, int SC, int DIR, int STEP>
__device__ __forceinline__ bool lane_walk(const float* __restrict__ src, float* __restrict__ dst, int nf, int nw,
                                          int nc, int64_t ws, int64_t cs, float h, int bf, int c) {
    const int lane = threadIdx.x;
    const int f = bf * 64 + lane;
    const bool live = f < nf;
    const int fl = live ? f : nf - 1;
    const int ofm = (fl > 0) ? -1 : 1;
    const int ofp = (fl < nf - 1) ? 1 : -1;
    const int64_t ocm = (c > 0) ? -cs : cs;
    const int64_t ocp = (c < nc - 1) ? cs : -cs;
    const SweepConst K = sweep_const(h);
    // byte offsets from the step's resource base (row - A), all >= 0
    const int64_t A = cs + ws + 1;
    const uint32_t v_fm = (uint32_t)((fl + ofm + A) * 4);
    const uint32_t v_fp = (uint32_t)((fl + ofp + A) * 4);
    const uint32_t v_cm = (uint32_t)((fl + ocm + A) * 4);
    const uint32_t v_cp = (uint32_t)((fl + ocp + A) * 4);
    const uint32_t v_nx = (uint32_t)((fl + (int64_t)DIR * ws + A) * 4);
    const uint32_t v_st = live ? (uint32_t)(f * 4) : 0x7ffffff0u;
    const uint32_t nrec = (uint32_t)((2 * A + nf + 64) * 4);
    const uint32_t nrow = (uint32_t)(nf * 4);
    const int64_t step = (int64_t)DIR * ws;
    int64_t row = (int64_t)c * cs + (int64_t)((DIR > 0) ? 0 : nw - 1) * ws;
    float behind = kFar;
    float cur = bload(rsrc_at(src + row - A, nrec), v_nx - (uint32_t)(step * 4));  // the first cell
    bool chg = false;
    int s0 = 0;
    // full groups: every step of the group has a next cell on the line
    for (; s0 + STEP < nw; s0 += STEP) {
        float fm[STEP], fp[STEP], cm[STEP], cp[STEP], nx[STEP];
#pragma unroll
        for (int k = 0; k < STEP; ++k) {
            const __amdgpu_buffer_rsrc_t r = rsrc_at(src + (row + k * step) - A, nrec);
            fm[k] = bload(r, v_fm);
            fp[k] = bload(r, v_fp);
            cm[k] = bload(r, v_cm);
            cp[k] = bload(r, v_cp);
            nx[k] = bload(r, v_nx);
        }
#pragma unroll
        for (int k = 0; k < STEP; ++k) {
            const float nv = lane_step<SF, SW, SC>(cur, fm[k], fp[k], cm[k], cp[k], behind, nx[k], K);
            chg |= nv < cur;
            bstore(rsrc_at(dst + row, nrow), v_st, nv);
            behind = nv;
            cur = nx[k];
            row += step;
        }
    }
    // the last steps, one at a time
    for (; s0 < nw; ++s0) {
        const __amdgpu_buffer_rsrc_t r = rsrc_at(src + row - A, nrec);
        const float a = bload(r, v_fm), b = bload(r, v_fp), cmv = bload(r, v_cm), cpv = bload(r, v_cp);
        const float next = (s0 + 1 < nw) ? bload(r, v_nx) : kFar;
        const float nv = lane_step<SF, SW, SC>(cur, a, b, cmv, cpv, behind, next, K);
        chg |= nv < cur;
        bstore(rsrc_at(dst + row, nrow), v_st, nv);
        behind = nv;
        cur = next;
        row += step;
    }
    return chg;
}

template <int SF, int SW, int SC, int DIR, int STEP, int ROWS, bool FMAJ>
__global__ __launch_bounds__(64 * ROWS) void sweep_lane_kernel(const float* __restrict__ src,
                                                               float* __restrict__ dst, int nf, int nw, int nc,
                                                               int64_t ws, int64_t cs, int nbf, unsigned nblocks,
                                                               float h, int* __restrict__ changed) {
    // FMAJ: the ROWS waves of a block take adjacent 64-cell segments of ONE
    // line row (long contiguous accesses); otherwise ROWS adjacent rows.
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    int bf, c;
    if (FMAJ) {
        const int nbg = (nbf + ROWS - 1) / ROWS;
        bf = __builtin_amdgcn_readfirstlane((int)(lb % (unsigned)nbg) * ROWS + (int)threadIdx.y);
        c = (int)(lb / (unsigned)nbg);
        if (bf >= nbf) return;
    } else {
        bf = (int)(lb % (unsigned)nbf);
        c = __builtin_amdgcn_readfirstlane((int)(lb / (unsigned)nbf) * ROWS + (int)threadIdx.y);
    }
    if (c >= nc) return;
    const bool chg = lane_walk<SF, SW, SC, DIR, STEP>(src, dst, nf, nw, nc, ws, cs, h, bf, c);
    if (__ballot(chg) != 0 && threadIdx.x == 0) atomicOr(changed, 1);
}

// The one-line walker over many lattices in one launch (4 rows per block, a
// job's blocks as sweep_lane_kernel's for it alone; J.nbf = 64-cell
// segments per row), gated and flagged per round like
// sweep_tile_batch_kernel.
template <int SF, int SW, int SC, int DIR, int STEP>
__global__ __launch_bounds__(256) void sweep_lane_batch_kernel(const SweepJob* __restrict__ jobs,
                                                               const unsigned* __restrict__ starts, int njobs,
                                                               unsigned nblocks, int r) {
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const int j = __builtin_amdgcn_readfirstlane(batch_job(starts, njobs, lb));
    const SweepJob J = jobs[j];
    if (r > 0 && J.flags[r - 1] == 0) return;
    const unsigned local = lb - starts[j];
    const int bf = (int)(local % (unsigned)J.nbf);
    const int c = __builtin_amdgcn_readfirstlane((int)(local / (unsigned)J.nbf) * 4 + (int)threadIdx.y);
    if (c >= J.nc) return;
    const bool chg = lane_walk<SF, SW, SC, DIR, STEP>(J.src, J.dst, J.nf, J.nw, J.nc, J.ws, J.cs, J.h, bf, c);
    if (__ballot(chg) != 0 && threadIdx.x == 0) atomicOr(J.flags + r, 1);
}

// ---- plane transpose (x-fastest <-> y-fastest) ----------------------------------
// Each z-plane is an (rows x cols) matrix; out[col][row] = in[row][col].
// 64x64 tiles through LDS (one padding column: conflict-free column reads).
// A block transposes kTrTiles tiles along a tile row: the next tile's loads
// are issued before the current one is stored (double-buffered LDS), which
// keeps the transpose at copy speed.
constexpr int kTrTiles = 2;

// With a sign volume (dense, laid out like `out`) every stored value gets the
// [F] apply on the way: the write-back of the rounds' y-fastest copy.
__device__ __forceinline__ void transpose_group(const float* __restrict__ in, float* __restrict__ out, int rows,
                                                int cols, int64_t in_pitch, int64_t in_plane, int64_t out_pitch,
                                                int64_t out_plane, int tiles_r, int tiles_c, unsigned lb,
                                                float (&tile)[2][64][65], const uint8_t* __restrict__ sign = nullptr,
                                                unsigned long long* __restrict__ far = nullptr) {
    const int groups_c = (tiles_c + kTrTiles - 1) / kTrTiles;
    const int per_plane = tiles_r * groups_c;
    const int z = (int)(lb / (unsigned)per_plane);
    const int t = (int)(lb % (unsigned)per_plane);
    const int tr = t / groups_c, tg = t % groups_c;
    const float* inz = in + (int64_t)z * in_plane;
    float* outz = out + (int64_t)z * out_plane;
    const uint8_t* signz = sign ? sign + (int64_t)z * out_plane : nullptr;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    const int r0 = tr * 64;
    float v[16];
    auto load = [&](int tc) {
        const int c = tc * 64 + tx;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int r = r0 + ty + 4 * k;
            v[k] = (r < rows && c < cols) ? inz[(int64_t)r * in_pitch + c] : 0.0f;
        }
    };
    const int tc0 = tg * kTrTiles, tc1 = min(tiles_c, tc0 + kTrTiles);
    unsigned nfar = 0;  // stored cells still at the sentinel (when `far` is given)
    load(tc0);
    for (int tc = tc0; tc < tc1; ++tc) {
        const int buf = (tc - tc0) & 1;
        const int r = r0 + tx;
        // this tile's sign bytes, requested before the LDS exchange so that
        // they arrive while it runs (not one round trip per store)
        uint8_t sg[16];
        if (signz) {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int c = tc * 64 + ty + 4 * k;
                sg[k] = (r < rows && c < cols) ? signz[(int64_t)c * out_pitch + r] : (uint8_t)0;
            }
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) tile[buf][ty + 4 * k][tx] = v[k];
        __syncthreads();  // also orders the previous use of the other buffer
        if (tc + 1 < tc1) load(tc + 1);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int c = tc * 64 + ty + 4 * k;
            if (r < rows && c < cols) {
                const int64_t o = (int64_t)c * out_pitch + r;
                const float v = tile[buf][tx][ty + 4 * k];
                if (far) nfar += v >= kFarTest;
                outz[o] = signz ? apply_sign(v, sg[k]) : v;
            }
        }
    }
    if (far) {
        for (int off = 32; off > 0; off >>= 1) nfar += __shfl_xor(nfar, off);
        if (__lane_id() == 0 && nfar) atomicAdd(far, (unsigned long long)nfar);
    }
}

__global__ __launch_bounds__(256) void transpose_planes_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                               int rows, int cols, int64_t in_pitch, int64_t in_plane,
                                                               int64_t out_pitch, int64_t out_plane, int tiles_r,
                                                               int tiles_c, unsigned nblocks) {
    __shared__ float tile[2][64][65];
    transpose_group(in, out, rows, cols, in_pitch, in_plane, out_pitch, out_plane, tiles_r, tiles_c,
                    xcd_logical_block(blockIdx.x, nblocks), tile);
}

// y-fastest copy -> the dense x-fastest field with the signs applied; with
// `far`, also the [D2] audit of the copy (cells >= 1e29 added to *far).
__global__ __launch_bounds__(256) void transpose_apply_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                              const uint8_t* __restrict__ sign, int rows, int cols,
                                                              int64_t in_pitch, int64_t in_plane, int tiles_r,
                                                              int tiles_c, unsigned nblocks,
                                                              unsigned long long* __restrict__ far) {
    __shared__ float tile[2][64][65];
    transpose_group(in, out, rows, cols, in_pitch, in_plane, rows, (int64_t)rows * cols, tiles_r, tiles_c,
                    xcd_logical_block(blockIdx.x, nblocks), tile, sign, far);
}

// The plane transposes of many lattices in one launch (the blocks of job j
// are those of transpose_planes_kernel for it alone), gated like the batched
// sweeps: round r of a lattice runs only if its round r - 1 changed a voxel.
__global__ __launch_bounds__(256) void transpose_batch_kernel(const TransJob* __restrict__ jobs,
                                                              const unsigned* __restrict__ starts, int njobs,
                                                              unsigned nblocks, int r) {
    __shared__ float tile[2][64][65];
    const unsigned lb = xcd_logical_block(blockIdx.x, nblocks);
    const int j = __builtin_amdgcn_readfirstlane(batch_job(starts, njobs, lb));
    const TransJob J = jobs[j];
    if (r > 0 && J.flags[r - 1] == 0) return;
    transpose_group(J.in, J.out, J.rows, J.cols, J.in_pitch, J.in_plane, J.out_pitch, J.out_plane, J.tiles_r,
                    J.tiles_c, lb - starts[j], tile);
}

// ---- classical 8-ordering fast sweeping (Zhao 2005) ---------------------------
// Ordering (sx, sy, sz) visits the cells lexicographically (z outermost, x
// innermost), each axis in its direction, updating in place: a cell reads its
// three upwind neighbours already updated and its three downwind ones not yet.
// Its diagonal levels kx + ky + kz = L (k = position along each axis in the
// ordering's direction) run one launch each: a cell of level L has its upwind
// neighbours on level L - 1 and its downwind ones on level L + 1, so it reads
// exactly what it reads in the lexicographic order, and the sweep equals the
// in-place lexicographic one bit for bit.  Cells of one level never read each
// other.  One thread per (kx, ky) of the level's box; kz follows.
__global__ void octant_level_kernel(float* __restrict__ g, int dx, int dy, int dz, float h, int sx, int sy, int sz,
                                    int level, int kx0, int nkx, int ky0, int64_t nk, int* __restrict__ changed) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool chg = false;
    if (t < nk) {
        const int kx = kx0 + (int)(t % nkx), ky = ky0 + (int)(t / nkx);
        const int kz = level - kx - ky;
        if (kz >= 0 && kz < dz) {
            const int x = sx > 0 ? kx : dx - 1 - kx;
            const int y = sy > 0 ? ky : dy - 1 - ky;
            const int z = sz > 0 ? kz : dz - 1 - kz;
            const int64_t ys = dx, zs = (int64_t)dx * dy;
            const int64_t i = (int64_t)z * zs + (int64_t)y * ys + x;
            const float S = 1e30f;
            const float cur = g[i];
            const float ax = fminf(x > 0 ? g[i - 1] : S, x < dx - 1 ? g[i + 1] : S);
            const float ay = fminf(y > 0 ? g[i - ys] : S, y < dy - 1 ? g[i + ys] : S);
            const float az = fminf(z > 0 ? g[i - zs] : S, z < dz - 1 ? g[i + zs] : S);
            const float u = godunov(ax, ay, az, h);
            if (u < cur) {
                g[i] = u;
                chg = true;
            }
        }
    }
    if (__ballot(chg) != 0 && (threadIdx.x & 63) == 0) atomicOr(changed, 1);
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

// Launch shape of the lane sweeps (steps per load group, waves per block);
// set by the tuning entry point, read at every launch.
// Class-free defaults: the shapes the autotuner picks for the 670M-voxel
// vessel lattice on MI355X (x 1 x 4, z 1 x 2 tiles, four-step load groups;
// the y-sweep of the y-fastest copy in the prefetching kernel, 4 waves of one
// plane each).
// Class-free defaults: the shape and load-group depth the autotuner keeps for
// the largest lattices (1 x 4 tiles for x and z on the y-fastest copy, two-step
// groups), so a caller that never tunes gets them where they matter most.
static SweepTuning g_sweep_tuning = {2, 4, 0, {14, 21, 14}, 16ll << 20, 0, {}, {}, {}, {}, 32, 1, 42};
SweepTuning sweep_tuning() { return g_sweep_tuning; }
bool set_sweep_tuning(int step, int rows, int fmajor) {
    if ((step != 2 && step != 4 && step != 8) || (rows != 4 && rows != 8 && rows != 16) ||
        (fmajor != 0 && fmajor != 1))
        return false;
    g_sweep_tuning.step = step;
    g_sweep_tuning.rows = rows;
    g_sweep_tuning.fmajor = fmajor;
    return true;
}
bool set_sweep_shape(int axis, int shape) {
    if (axis < 0 || axis > 2 || (shape != 1 && shape != 21 && shape != 22 && shape != 12 && shape != 14)) return false;
    g_sweep_tuning.shape[axis] = shape;
    return true;
}
int shape_class(int64_t voxels) {
    if (voxels <= (int64_t)1 << 21) return 0;
    if (voxels <= (int64_t)1 << 24) return 1;
    if (voxels <= (int64_t)1 << 27) return 2;
    return 3;
}
int effective_shape(int axis, int64_t voxels) {
    const int v = g_sweep_tuning.cls_shape[shape_class(voxels)][axis];
    return v ? v : g_sweep_tuning.shape[axis];
}
int effective_step(int64_t voxels) {
    const int v = g_sweep_tuning.cls_step[shape_class(voxels)];
    return v ? v : g_sweep_tuning.step;
}
bool set_run_chunk(int cells) {
    if (cells != 16 && cells != 32) return false;
    g_sweep_tuning.run_chunk = cells;
    return true;
}
bool set_run_block(int shape) {
    if (shape != 0 && shape != 41 && shape != 42) return false;
    g_sweep_tuning.run_block = shape;
    return true;
}
bool set_fsm_layout(int layout) {
    if (layout != 0 && layout != 1) return false;
    g_sweep_tuning.fsm_layout = layout;
    return true;
}
bool set_class_shape(int cls, int axis, int shape) {
    if (cls < 0 || cls >= kShapeClasses || axis < 0 || axis > 2 ||
        (shape != 0 && shape != 1 && shape != 21 && shape != 22 && shape != 12 && shape != 14))
        return false;
    g_sweep_tuning.cls_shape[cls][axis] = shape;
    return true;
}
bool set_class_step(int cls, int step) {
    if (cls < 0 || cls >= kShapeClasses || (step != 0 && step != 2 && step != 4 && step != 8)) return false;
    g_sweep_tuning.cls_step[cls] = step;
    return true;
}
bool set_batch_class_shape(int cls, int axis, int shape) {
    if (cls < 0 || cls >= kShapeClasses || axis < 0 || axis > 2 ||
        (shape != 0 && shape != 1 && shape != 21 && shape != 22 && shape != 12 && shape != 14))
        return false;
    g_sweep_tuning.bcls_shape[cls][axis] = shape;
    return true;
}
bool set_batch_class_step(int cls, int step) {
    if (cls < 0 || cls >= kShapeClasses || (step != 0 && step != 2 && step != 4 && step != 8)) return false;
    g_sweep_tuning.bcls_step[cls] = step;
    return true;
}
void set_fsm_batching(int64_t max_voxels, int graph) {
    g_sweep_tuning.batch_rounds_max_voxels = max_voxels;
    g_sweep_tuning.fsm_graph = graph ? 1 : 0;
}

namespace {

template <int SF, int SW, int SC, int DIR, int STEP, int ROWS>
hipError_t launch_lane_v(const float* src, float* dst, int nf, int nw, int nc, int64_t ws, int64_t cs, float h,
                         int* changed, hipStream_t s) {
    const int nbf = (nf + 63) / 64;
    dim3 blk(64, ROWS);
    if (sweep_tuning().fmajor) {
        const int64_t blocks = (int64_t)((nbf + ROWS - 1) / ROWS) * nc;
        sweep_lane_kernel<SF, SW, SC, DIR, STEP, ROWS, true><<<(unsigned)blocks, blk, 0, s>>>(
            src, dst, nf, nw, nc, ws, cs, nbf, (unsigned)blocks, h, changed);
    } else {
        const int64_t blocks = (int64_t)nbf * ((nc + ROWS - 1) / ROWS);
        sweep_lane_kernel<SF, SW, SC, DIR, STEP, ROWS, false><<<(unsigned)blocks, blk, 0, s>>>(
            src, dst, nf, nw, nc, ws, cs, nbf, (unsigned)blocks, h, changed);
    }
    return hipGetLastError();
}

template <int SF, int SW, int SC, int DIR, int STEP, int NF, int NC>
hipError_t launch_tile(const float* src, float* dst, int nf, int nw, int nc, int64_t ws, int64_t cs, float h,
                       int* changed, hipStream_t s) {
    const int nbf = (nf + 64 * NF - 1) / (64 * NF);
    const int ngroups = (nc + NC - 1) / NC;
    const int64_t blocks = (int64_t)nbf * ((ngroups + 3) / 4);
    sweep_tile_kernel<SF, SW, SC, DIR, STEP, NF, NC><<<(unsigned)blocks, dim3(64, 4), 0, s>>>(
        src, dst, nf, nw, nc, ws, cs, nbf, (unsigned)blocks, h, changed);
    return hipGetLastError();
}

template <int SF, int SW, int SC, int DIR>
hipError_t launch_lane_d(const float* src, float* dst, int nf, int nw, int nc, int64_t ws, int64_t cs, float h,
                         int* changed, hipStream_t s) {
    const SweepTuning t = sweep_tuning();
    // the walk axis names the orientation (the x-sweep runs on the y-fastest
    // copy); shape and step are those of the lattice's size class
    const int64_t voxels = (int64_t)nf * nw * nc;
    const int shape = effective_shape(SW, voxels);
    const int step = effective_step(voxels);
#define RHZ_TILE(NF, NC)                                                                                    \
    if (shape == 10 * NF + NC) {                                                                             \
        if (step == 2) return launch_tile<SF, SW, SC, DIR, 2, NF, NC>(src, dst, nf, nw, nc, ws, cs, h, changed, s); \
        if (step == 8) return launch_tile<SF, SW, SC, DIR, 8, NF, NC>(src, dst, nf, nw, nc, ws, cs, h, changed, s); \
        return launch_tile<SF, SW, SC, DIR, 4, NF, NC>(src, dst, nf, nw, nc, ws, cs, h, changed, s);         \
    }
    RHZ_TILE(2, 1)
    RHZ_TILE(2, 2)
    RHZ_TILE(1, 2)
    RHZ_TILE(1, 4)
#undef RHZ_TILE
    if (step == 8) {
        if (t.rows == 4) return launch_lane_v<SF, SW, SC, DIR, 8, 4>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
        if (t.rows == 16) return launch_lane_v<SF, SW, SC, DIR, 8, 16>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
        return launch_lane_v<SF, SW, SC, DIR, 8, 8>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
    }
    if (step == 2) {
        if (t.rows == 4) return launch_lane_v<SF, SW, SC, DIR, 2, 4>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
        if (t.rows == 16) return launch_lane_v<SF, SW, SC, DIR, 2, 16>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
        return launch_lane_v<SF, SW, SC, DIR, 2, 8>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
    }
    if (t.rows == 4) return launch_lane_v<SF, SW, SC, DIR, 4, 4>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
    if (t.rows == 16) return launch_lane_v<SF, SW, SC, DIR, 4, 16>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
    return launch_lane_v<SF, SW, SC, DIR, 4, 8>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
}

template <int SF, int SW, int SC>
hipError_t launch_lane(const float* src, float* dst, int nf, int nw, int nc, int64_t ws, int64_t cs, float h, int dir,
                       int* changed, hipStream_t s) {
    if (dir > 0) return launch_lane_d<SF, SW, SC, 1>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
    return launch_lane_d<SF, SW, SC, -1>(src, dst, nf, nw, nc, ws, cs, h, changed, s);
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
    // 32-bit byte offsets: the largest buffer record, 4 (4 dx dy + 3 max(dx, dy)
    // + 130) bytes, stays below 2^32 for planes of < 2^27 cells.  The tile
    // sweep's short square root needs 2^-30 <= h <= 2^30.
    const int64_t dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    return dx >= 2 && dy >= 2 && dz >= 2 && dx * dy < (int64_t(1) << 27) && g.h >= 0x1.0p-30f &&
           g.h <= 0x1.0p30f;
}

hipError_t launch_transpose_xy(const float* in, float* out, const Grid& g, bool to_yx, hipStream_t s,
                               const FieldLayout& lin, const FieldLayout& lout) {
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    const int rows = to_yx ? dy : dx, cols = to_yx ? dx : dy;
    if (lin.pitch < cols || lout.pitch < rows || lin.plane < lin.pitch * rows || lout.plane < lout.pitch * cols)
        return hipErrorInvalidValue;
    const int tr = (rows + 63) / 64, tc = (cols + 63) / 64;
    const int64_t blocks = (int64_t)tr * ((tc + kTrTiles - 1) / kTrTiles) * dz;
    transpose_planes_kernel<<<(unsigned)blocks, 256, 0, s>>>(in, out, rows, cols, lin.pitch, lin.plane, lout.pitch,
                                                            lout.plane, tr, tc, (unsigned)blocks);
    return hipGetLastError();
}

hipError_t launch_transpose_apply(const float* in, const FieldLayout& lin, float* out, const uint8_t* sign,
                                  const Grid& g, hipStream_t s, unsigned long long* far) {
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    const int rows = dx, cols = dy;  // in: rows along y (y-fastest); out: dense x-fastest
    if (lin.pitch < cols || lin.plane < lin.pitch * rows) return hipErrorInvalidValue;
    const int tr = (rows + 63) / 64, tc = (cols + 63) / 64;
    const int64_t blocks = (int64_t)tr * ((tc + kTrTiles - 1) / kTrTiles) * dz;
    transpose_apply_kernel<<<(unsigned)blocks, 256, 0, s>>>(in, out, sign, rows, cols, lin.pitch, lin.plane, tr, tc,
                                                           (unsigned)blocks, far);
    return hipGetLastError();
}

namespace {
// Every binary32 bit pattern in [lo, hi): the sweeps' square root against IEEE
// sqrt (binary32, and binary64 rounded to binary32: the square root of a
// binary32 value is never close enough to a binary32 midpoint for that double
// rounding to differ); mismatches are added to *bad.
__global__ __launch_bounds__(256) void sqrt_selftest_kernel(uint32_t lo, uint32_t hi,
                                                            unsigned long long* __restrict__ bad) {
    const uint64_t n = (uint64_t)(hi - lo);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float x = __uint_as_float(lo + (uint32_t)i);
        const uint32_t r = __float_as_uint(sqrt_rn_normal(x));
        c += (r != __float_as_uint(sqrtf(x))) + (r != __float_as_uint((float)sqrt((double)x)));
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (__lane_id() == 0 && c) atomicAdd(bad, (unsigned long long)c);
}
}  // namespace

hipError_t launch_sqrt_selftest(unsigned long long* bad, hipStream_t s) {
    // [2^-96, +inf): the range every sweep square root argument lies in
    sqrt_selftest_kernel<<<8192, 256, 0, s>>>(0x0f800000u, 0x7f800000u, bad);
    return hipGetLastError();
}

hipError_t launch_transpose_xy(const float* in, float* out, const Grid& g, bool to_yx, hipStream_t s) {
    const FieldLayout xy = dense_xy(g), yx = dense_yx(g);
    return launch_transpose_xy(in, out, g, to_yx, s, to_yx ? xy : yx, to_yx ? yx : xy);
}

hipError_t launch_sweep_yx(const float* src, float* dst, const Grid& g, int dir, int* changed, hipStream_t s,
                           const FieldLayout& L) {
    // y-fastest layout: lanes over y (Godunov slot 1), walk x (slot 0, stride = row pitch), cross z (slot 2, plane)
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    return launch_lane<1, 0, 2>(src, dst, dy, dx, dz, L.pitch, L.plane, g.h, dir, changed, s);
}

hipError_t launch_sweep_yx(const float* src, float* dst, const Grid& g, int dir, int* changed, hipStream_t s) {
    return launch_sweep_yx(src, dst, g, dir, changed, s, dense_yx(g));
}

namespace {
template <int DIR, int CH>
hipError_t launch_run(const float* src, float* dst, const Grid& g, const FieldLayout& L, int* changed, hipStream_t s) {
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    const int nbx = (nx + 63) / 64;
    const int64_t blocks = (int64_t)nbx * ((nz + 3) / 4);
    sweep_run_kernel<DIR, CH><<<(unsigned)blocks, dim3(64, 4), 0, s>>>(src, dst, nx, ny, nz, L.pitch, L.plane, nbx,
                                                                       (unsigned)blocks, g.h, changed);
    return hipGetLastError();
}
template <int DIR, int W, int CH, int C, bool ZMAJ>
hipError_t launch_runp(const float* src, float* dst, const Grid& g, const FieldLayout& L, int* changed,
                       hipStream_t s) {
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    const int nbx = (nx + 61) / 62;
    const int64_t blocks = (int64_t)nbx * ((nz + W * C - 1) / (W * C));
    sweep_runp_kernel<DIR, CH, W, C, ZMAJ><<<(unsigned)blocks, dim3(64, W), 0, s>>>(
        src, dst, nx, ny, nz, L.pitch, L.plane, nbx, (unsigned)blocks, g.h, changed);
    return hipGetLastError();
}
// "sweep_run_block" 41: 4 waves of one plane each, 16-cell chunks; 42: the
// same with plane-group-major block order
template <int DIR>
hipError_t launch_runp(int shape, const float* src, float* dst, const Grid& g, const FieldLayout& L, int* changed,
                       hipStream_t s) {
    if (shape == 42) return launch_runp<DIR, 4, 16, 1, true>(src, dst, g, L, changed, s);
    return launch_runp<DIR, 4, 16, 1, false>(src, dst, g, L, changed, s);
}
}  // namespace

hipError_t launch_sweep_on_yx(const float* src, float* dst, const Grid& g, int axis, int dir, int* changed,
                              hipStream_t s, const FieldLayout& L) {
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    if (axis == 0) return launch_sweep_yx(src, dst, g, dir, changed, s, L);
    if (axis == 2)  // lanes y (slot 1, contiguous), walk z (slot 2, plane), cross x (slot 0, row pitch)
        return launch_lane<1, 2, 0>(src, dst, dy, dz, dx, L.plane, L.pitch, g.h, dir, changed, s);
    // axis 1: the walk runs along the rows, which must be padded to 32 floats
    if (L.pitch % 32 != 0 || !lane_sweeps_ok(g)) return hipErrorInvalidValue;
    const int wb = sweep_tuning().run_block;
    if (wb != 0)  // the prefetching kernel
        return dir > 0 ? launch_runp<1>(wb, src, dst, g, L, changed, s) : launch_runp<-1>(wb, src, dst, g, L, changed, s);
    if (sweep_tuning().run_chunk == 16)
        return dir > 0 ? launch_run<1, 16>(src, dst, g, L, changed, s) : launch_run<-1, 16>(src, dst, g, L, changed, s);
    return dir > 0 ? launch_run<1, 32>(src, dst, g, L, changed, s) : launch_run<-1, 32>(src, dst, g, L, changed, s);
}

hipError_t launch_sweep(const float* src, float* dst, const Grid& g, int axis, int dir, int* changed,
                        hipStream_t s, const FieldLayout& L) {
    const int dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    if (axis == 0 || !lane_sweeps_ok(g)) {
        const FieldLayout d = dense_xy(g);
        if (L.pitch != d.pitch || L.plane != d.plane) return hipErrorInvalidValue;  // generic kernels: dense only
        return launch_generic_sweep(src, dst, g, axis, dir, changed, s);
    }
    if (axis == 1)  // lanes x (slot 0), walk y (slot 1, stride = row pitch), cross z (slot 2, plane)
        return launch_lane<0, 1, 2>(src, dst, dx, dy, dz, L.pitch, L.plane, g.h, dir, changed, s);
    // axis 2: lanes x (slot 0), walk z (slot 2, stride = plane), cross y (slot 1, stride = row pitch)
    return launch_lane<0, 2, 1>(src, dst, dx, dz, dy, L.plane, L.pitch, g.h, dir, changed, s);
}

hipError_t launch_sweep(const float* src, float* dst, const Grid& g, int axis, int dir, int* changed,
                        hipStream_t s) {
    return launch_sweep(src, dst, g, axis, dir, changed, s, dense_xy(g));
}

// The FSM's working layouts: rows rounded up to 32 floats (128 B), so that
// every 64-cell segment of a row spans two cache lines instead of three;
// planes hold the lattice's own row count (rounding it up as well measured
// the same).  Dense when the padded planes would leave the lane sweeps'
// 32-bit offset range.
void fsm_layouts(const Grid& g, FieldLayout* xy, FieldLayout* yx) {
    const int64_t dx = g.dims[0], dy = g.dims[1];
    const int64_t px = (dx + 31) / 32 * 32, py = (dy + 31) / 32 * 32;
    const bool ok = lane_sweeps_ok(g) && px * dy < (int64_t(1) << 27) && py * dx < (int64_t(1) << 27);
    *xy = ok ? FieldLayout{px, px * dy} : dense_xy(g);
    *yx = ok ? FieldLayout{py, py * dx} : dense_yx(g);
}

// Tile shape (NF x NC) of the batched sweeps per orientation and size class:
// the batched stage's own per-class shape for that walk axis, else the
// class-free one; NC = 0 for the one-line walker (sweep_lane_batch_kernel,
// 4 rows per block).
static void batch_shape(int orient, int cls, int* nf_tile, int* nc_tile) {
    const int walk = orient == 0 ? 0 : orient;  // walk axis: x, y, z
    const SweepTuning t = sweep_tuning();
    const int shape = t.bcls_shape[cls][walk] ? t.bcls_shape[cls][walk] : t.shape[walk];
    if (shape == 1) {  // one-line walker: one 64-cell segment, one row per wave
        *nf_tile = 1;
        *nc_tile = 0;
        return;
    }
    *nf_tile = shape / 10;
    *nc_tile = shape % 10;
}

int64_t sweep_batch_blocks(int orient, int cls, int nf, int nc, int* nbf) {
    int NF = 2, NC = 1;
    batch_shape(orient, cls, &NF, &NC);
    *nbf = (nf + 64 * NF - 1) / (64 * NF);
    const int ngroups = NC ? (nc + NC - 1) / NC : nc;
    return (int64_t)*nbf * ((ngroups + 3) / 4);
}

namespace {
template <int SF, int SW, int SC, int DIR, int STEP>
hipError_t launch_sweep_batch_s(int NF, int NC, const SweepJob* jobs, const unsigned* starts, int njobs,
                                unsigned nblocks, int r, hipStream_t s) {
    if (NC == 0) {
        sweep_lane_batch_kernel<SF, SW, SC, DIR, STEP><<<nblocks, dim3(64, 4), 0, s>>>(jobs, starts, njobs, nblocks, r);
        return hipGetLastError();
    }
#define RHZ_BTILE(A, B)                                                                                  \
    if (NF == A && NC == B) {                                                                            \
        sweep_tile_batch_kernel<SF, SW, SC, DIR, STEP, A, B><<<nblocks, dim3(64, 4), 0, s>>>(jobs, starts, njobs, \
                                                                                           nblocks, r);   \
        return hipGetLastError();                                                                        \
    }
    RHZ_BTILE(2, 1)
    RHZ_BTILE(2, 2)
    RHZ_BTILE(1, 2)
    RHZ_BTILE(1, 4)
#undef RHZ_BTILE
    return hipErrorInvalidValue;
}

template <int SF, int SW, int SC, int DIR>
hipError_t launch_sweep_batch_d(int NF, int NC, int step, const SweepJob* jobs, const unsigned* starts, int njobs,
                                unsigned nblocks, int r, hipStream_t s) {
    if (step == 4) return launch_sweep_batch_s<SF, SW, SC, DIR, 4>(NF, NC, jobs, starts, njobs, nblocks, r, s);
    if (step == 8) return launch_sweep_batch_s<SF, SW, SC, DIR, 8>(NF, NC, jobs, starts, njobs, nblocks, r, s);
    return launch_sweep_batch_s<SF, SW, SC, DIR, 2>(NF, NC, jobs, starts, njobs, nblocks, r, s);
}
}  // namespace

hipError_t launch_sweep_batch(int orient, int cls, int dir, const SweepJob* jobs, const unsigned* starts, int njobs,
                              unsigned nblocks, int r, hipStream_t s) {
    if (njobs <= 0 || nblocks == 0) return hipSuccess;
    int NF = 2, NC = 1;
    batch_shape(orient, cls, &NF, &NC);
    const SweepTuning t = sweep_tuning();
    const int st = t.bcls_step[cls] ? t.bcls_step[cls] : t.step;
    if (orient == 0)  // y-fastest copy: lanes y (slot 1), walk x (slot 0), cross z (slot 2)
        return dir > 0 ? launch_sweep_batch_d<1, 0, 2, 1>(NF, NC, st, jobs, starts, njobs, nblocks, r, s)
                       : launch_sweep_batch_d<1, 0, 2, -1>(NF, NC, st, jobs, starts, njobs, nblocks, r, s);
    if (orient == 1)  // lanes x, walk y, cross z
        return dir > 0 ? launch_sweep_batch_d<0, 1, 2, 1>(NF, NC, st, jobs, starts, njobs, nblocks, r, s)
                       : launch_sweep_batch_d<0, 1, 2, -1>(NF, NC, st, jobs, starts, njobs, nblocks, r, s);
    // lanes x, walk z, cross y
    return dir > 0 ? launch_sweep_batch_d<0, 2, 1, 1>(NF, NC, st, jobs, starts, njobs, nblocks, r, s)
                   : launch_sweep_batch_d<0, 2, 1, -1>(NF, NC, st, jobs, starts, njobs, nblocks, r, s);
}

int64_t transpose_batch_blocks(int rows, int cols, int planes, int* tiles_r, int* tiles_c) {
    *tiles_r = (rows + 63) / 64;
    *tiles_c = (cols + 63) / 64;
    return (int64_t)*tiles_r * ((*tiles_c + kTrTiles - 1) / kTrTiles) * planes;
}

hipError_t launch_transpose_batch(const TransJob* jobs, const unsigned* starts, int njobs, unsigned nblocks, int r,
                                  hipStream_t s) {
    if (njobs <= 0 || nblocks == 0) return hipSuccess;
    transpose_batch_kernel<<<nblocks, 256, 0, s>>>(jobs, starts, njobs, nblocks, r);
    return hipGetLastError();
}

hipError_t launch_sweep_octant(float* g, const Grid& gr, int sx, int sy, int sz, int* changed, hipStream_t s) {
    const int dx = gr.dims[0], dy = gr.dims[1], dz = gr.dims[2];
    if (dx <= 0 || dy <= 0 || dz <= 0) return hipSuccess;
    const int levels = dx + dy + dz - 2;
    for (int L = 0; L < levels; ++L) {
        const int kx0 = std::max(0, L - (dy - 1) - (dz - 1)), kx1 = std::min(dx - 1, L);
        const int ky0 = std::max(0, L - (dx - 1) - (dz - 1)), ky1 = std::min(dy - 1, L);
        if (kx1 < kx0 || ky1 < ky0) continue;
        const int nkx = kx1 - kx0 + 1;
        const int64_t nk = (int64_t)nkx * (ky1 - ky0 + 1);
        octant_level_kernel<<<(unsigned)((nk + 255) / 256), 256, 0, s>>>(g, dx, dy, dz, gr.h, sx, sy, sz, L, kx0, nkx,
                                                                          ky0, nk, changed);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_redblack(float* g, const int32_t dims[3], float h, int parity, hipStream_t s) {
    int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    redblack_kernel<<<(unsigned)b, 256, 0, s>>>(g, dims[0], dims[1], dims[2], h, parity);
    return hipGetLastError();
}

}  // namespace rhz
