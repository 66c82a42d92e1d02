// k_learned.hip -- risk-ranked exact re-probing (the learned sign gate).
//
// After the band splat and the scanline sign pass, the only voxels whose
// scanline sign can disagree with the exact winding sign on a closed mesh are
// (a) voxels within the jitter offset of the surface (the ray runs
// (0.037, 0.051) h away from the voxel centre) and (b) voxels of deferred
// columns, which the exact fixup already decided.  These kernels:
//   * compact the candidate voxels (band distance <= r h),
//   * compute 12 geometric features per candidate from the column's
//     candidate-triangle list,
//   * rank them (the proposer network or the planar-margin rule) and select
//     the top-ranked ones with a score histogram,
//   * commit exact decisions, and grow the set around every mismatch found
//     (26-neighbourhood, each voxel probed at most once: a bitmap).
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int per = kBlock) {
    int64_t b = (n + per - 1) / per;
    if (b > 8192) b = 8192;
    return (unsigned)(b < 1 ? 1 : b);
}

// Candidate compaction: a block handles chunks of 256 x 16 float4 (16,384
// voxels), reads them coalesced, ranks its hits with a block-wide prefix sum
// and reserves output slots with ONE atomic per chunk (a per-wave atomic on a
// single counter serialises at the L2 for a sparse but surface-wide list).
constexpr int kCandBlock = 256;
constexpr int kCandPer = 16;  // float4 per thread per chunk

__global__ __launch_bounds__(kCandBlock) void candidates_kernel(const float* __restrict__ d, int64_t n, float thr,
                                                                int64_t* __restrict__ list, int64_t cap,
                                                                unsigned long long* counter) {
    __shared__ unsigned wave_tot[kCandBlock / 64];
    __shared__ unsigned long long chunk_base;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t n4 = n >> 2;
    const float4* d4 = reinterpret_cast<const float4*>(d);
    const int64_t per_chunk = (int64_t)kCandBlock * kCandPer;
    const int64_t nchunks = (n4 + per_chunk - 1) / per_chunk;
    for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int64_t base4 = ch * per_chunk;
        unsigned long long m = 0;
#pragma unroll
        for (int i = 0; i < kCandPer; ++i) {
            const int64_t q = base4 + (int64_t)i * kCandBlock + tid;
            if (q < n4) {
                const float4 v = d4[q];
                m |= (unsigned long long)((v.x <= thr) | ((v.y <= thr) << 1) | ((v.z <= thr) << 2) |
                                          ((v.w <= thr) << 3))
                     << (4 * i);
            }
        }
        const unsigned cnt = (unsigned)__popcll(m);
        // inclusive wave scan
        unsigned incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wave_tot[wv] = incl;
        __syncthreads();
        unsigned before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kCandBlock / 64; ++w) {
            const unsigned t = wave_tot[w];
            before += (w < wv) ? t : 0u;
            total += t;
        }
        if (tid == 0) chunk_base = total ? atomicAdd(counter, (unsigned long long)total) : 0ull;
        __syncthreads();
        unsigned long long pos = chunk_base + before + incl - cnt;
        while (m) {
            const int b = __ffsll((long long)m) - 1;
            m &= m - 1;
            const int i = b >> 2, j = b & 3;
            if (pos < (unsigned long long)cap) list[pos] = (base4 + (int64_t)i * kCandBlock + tid) * 4 + j;
            ++pos;
        }
        __syncthreads();  // wave_tot / chunk_base are reused by the next chunk
    }
    // the n % 4 trailing voxels
    if (blockIdx.x == 0 && tid < 64) {
        const int64_t t = (n4 << 2) + tid;
        const bool want = (t < n) && (d[t] <= thr);
        long long slot = wave_append(counter, want);
        if (want && slot < cap) list[slot] = t;
    }
}

// Does the horizontal segment from the voxel centre P0 to its jittered ray
// point P1 = P0 + (jx, jy, 0) h possibly cross triangle t?  The scanline
// pass classified P1; the winding number of P0 equals that of P1 unless the
// segment crosses the surface.  Conservative: "no" only when both ends lie
// beyond eps on one side of t's plane, or the segment meets the plane more
// than eps away from t.  Zero-area triangles (zero normal) bound no volume.
__device__ __forceinline__ bool segment_may_cross(float px, float py, float pz, float rx, float ry, const Tri9& t,
                                                  float nx, float ny, float nz, float eps, float half_len) {
    if (nx == 0.0f && ny == 0.0f && nz == 0.0f) return false;
    const float s0 = nx * (px - t.ax) + ny * (py - t.ay) + nz * (pz - t.az);
    const float s1 = nx * (rx - t.ax) + ny * (ry - t.ay) + nz * (pz - t.az);
    if ((s0 > eps && s1 > eps) || (s0 < -eps && s1 < -eps)) return false;
    const bool flat = fabsf(s0) <= eps && fabsf(s1) <= eps;  // segment (nearly) in the plane
    float lam = 0.5f;
    if (!flat) {
        const float den = s0 - s1;
        lam = (fabsf(den) > 0.0f) ? fminf(fmaxf(s0 / den, 0.0f), 1.0f) : 0.5f;
    }
    const float qx = px + lam * (rx - px), qy = py + lam * (ry - py);
    const float r = flat ? (half_len + eps) : eps;
    return point_tri_dist2(qx, qy, pz, t) <= r * r;
}

// Features of candidate voxel v (nearest triangles searched in the voxel's
// scanline supercell list), and risk[i] = 1 when the segment certificate
// above cannot rule out a scanline sign error:
//  0 d/h  1 scan sign (the z ray)  2-4 nearest-triangle normal
//  5 inside by the signed crossings of an x ray (axis_gather_kernel)
//  6 planar margin  7 inside by a y ray (axis_gather_kernel)
//  8 log1p(#triangles within 1 h)  9 normal spread within 1.5 h
//  10 the certificate flag itself  11 distance from the voxel centre to the
//  nearest crossing of its jittered column ray (voxels)
// z extent of every triangle (the features kernel's cheap reject).
// The largest z extent goes to *max_ext (atomic max on the bit pattern of a
// non-negative float).
__global__ void tri_zrange_kernel(const float* __restrict__ tris, int64_t nt, float2* __restrict__ zr,
                                  unsigned* __restrict__ max_ext) {
    float ext = 0.0f;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
        const Tri9 v = load_tri(tris, t);
        const float lo = fminf(v.az, fminf(v.bz, v.cz)), hi = fmaxf(v.az, fmaxf(v.bz, v.cz));
        zr[t] = make_float2(lo, hi);
        ext = fmaxf(ext, hi - lo);
    }
    atomicMax(max_ext, __float_as_uint(ext));
}

// Order-preserving map of a float to an unsigned key (all non-NaN values).
__device__ __forceinline__ uint32_t zkey_of(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// key[e] = (zkey(zmin of ids[e]) << 32) | ids[e]: sorting a supercell's keys
// orders its list by lowest z, then triangle id.
__global__ void zsort_keys_kernel(const int32_t* __restrict__ ids, int64_t n, const float2* __restrict__ zr,
                                  unsigned long long* __restrict__ key) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int32_t t = ids[e];
    key[e] = ((unsigned long long)zkey_of(zr[t].x) << 32) | (uint32_t)t;
}

// Every feature of this kernel depends only on triangles within 4 h of the
// voxel (ray crossings are capped at 2 h, the rest are nearer; the nearest
// distance of a candidate is below the candidate radius), so a triangle whose
// z extent is farther than 4.004 h from the voxel is skipped before its
// distance is computed.  The column list is
// sorted by (lowest z, triangle id), so the walk starts at the first entry
// whose lowest z reaches pz - reach - (largest z extent) and stops after the
// last whose lowest z is within pz + reach, both widened by a relative margin
// far above float rounding: every entry left out would fail the exact reach
// test, so the same triangles pass it as with the whole list, and they are
// taken in list order (which fixes the first of tied nearest triangles and
// the order of the normal sums).  The window is walked kFeatBatch entries at
// a time, ids and z extents loaded together (the walk is latency-bound).
constexpr int kFeatBatch = 8;  // column-list entries whose loads are issued together

__global__ void features_kernel(const float* __restrict__ tris, const float4* __restrict__ xyab,
                                const float2* __restrict__ zr, const int64_t* __restrict__ start,
                                const unsigned long long* __restrict__ keys, const unsigned* __restrict__ max_ext,
                                const uint8_t* __restrict__ sign, const int64_t* __restrict__ list, int64_t n,
                                int mnx, int mny, int mnz, int dx, int dy, float h, int sup, int sdx, float jx,
                                float jy, float* __restrict__ feats, uint8_t* __restrict__ risk,
                                unsigned long long* __restrict__ n_risky) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t v = list[i];
    const int x = (int)(v % dx);
    const int64_t r = v / dx;
    const int y = (int)(r % dy);
    const int z = (int)(r / dy);
    const float px = lattice(mnx, x, h), py = lattice(mny, y, h), pz = lattice(mnz, z, h);
    const float rx = ((float)(mnx + x) + jx) * h;
    const float ry = ((float)(mny + y) + jy) * h;
    const float h2 = h * h;
    float best = 3.0e38f;
    int best_t = -1;
    int near_cnt = 0;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    int spread_cnt = 0;
    float zhit = 2.0f * h;
    bool risky = false;
    const float eps = 2e-3f * h;
    const float half_len = 0.5f * sqrtf(jx * jx + jy * jy) * h;
    const int64_t s = (int64_t)(y / sup) * sdx + (x / sup);
    const float reach = 4.004f * h;
    const float ext = __uint_as_float(*max_ext);
    const float margin = 0x1.0p-10f * (fabsf(pz) + reach + ext);
    const uint32_t klo = zkey_of(pz - reach - ext - margin), khi = zkey_of(pz + reach + margin);
    int64_t l0 = start[s], l1 = start[s + 1];
    {
        // first entry with lowest-z key >= klo
        int64_t a = l0, b = l1;
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            if ((uint32_t)(keys[m] >> 32) < klo) a = m + 1;
            else b = m;
        }
        l0 = a;
        // first entry with lowest-z key > khi
        b = l1;
        while (a < b) {
            const int64_t m = (a + b) >> 1;
            if ((uint32_t)(keys[m] >> 32) <= khi) a = m + 1;
            else b = m;
        }
        l1 = a;
    }
    for (int64_t eb = l0; eb < l1; eb += kFeatBatch) {
        // every index is clamped into [l0, l1), so the loads need no guard
        int tid[kFeatBatch];
#pragma unroll
        for (int j = 0; j < kFeatBatch; ++j) tid[j] = (int)(uint32_t)keys[min(eb + j, l1 - 1)];
        bool in_reach[kFeatBatch];
#pragma unroll
        for (int j = 0; j < kFeatBatch; ++j) {
            const float2 q = zr[tid[j]];
            in_reach[j] = (eb + j < l1) && !(pz < q.x - reach || pz > q.y + reach);
        }
#pragma unroll
        for (int j = 0; j < kFeatBatch; ++j) {
            if (!in_reach[j]) continue;
            const int ti = tid[j];
            const Tri9 t = load_tri(tris, ti);
            const float d2 = point_tri_dist2(px, py, pz, t);
            if (!risky && d2 <= (2.0f * half_len + eps) * (2.0f * half_len + eps)) {
                const float4 tn = reinterpret_cast<const float4*>(tris + 12 * (int64_t)ti)[2];
                risky = segment_may_cross(px, py, pz, rx, ry, t, tn.y, tn.z, tn.w, eps, half_len);
            }
            if (d2 < best) {
                best = d2;
                best_t = ti;
            }
            if (d2 <= h2) ++near_cnt;
            if (d2 <= 2.25f * h2) {
                const float4 nn = reinterpret_cast<const float4*>(tris + 12 * (int64_t)ti)[2];
                sx += nn.y;
                sy += nn.z;
                sz += nn.w;
                ++spread_cnt;
            }
            // crossing of the jittered column ray (same tests as the scan)
            const float4 ab = xyab[ti];
            if (!(rx < ab.x || rx > ab.z || ry < ab.y || ry > ab.w)) {
                float e0 = (t.bx - t.ax) * (ry - t.ay) - (t.by - t.ay) * (rx - t.ax);
                float e1 = (t.cx - t.bx) * (ry - t.by) - (t.cy - t.by) * (rx - t.bx);
                float e2 = (t.ax - t.cx) * (ry - t.cy) - (t.ay - t.cy) * (rx - t.cx);
                bool pos = (e0 > 0.0f && e1 > 0.0f && e2 > 0.0f);
                bool neg = (e0 < 0.0f && e1 < 0.0f && e2 < 0.0f);
                float area2 = e0 + e1 + e2;
                if ((pos || neg) && fabsf(area2) >= 1e-12f) {
                    float zh = (e1 * t.az + e2 * t.bz + e0 * t.cz) / area2;
                    zhit = fminf(zhit, fabsf(zh - pz));
                }
            }
        }
    }
    float nx = 0.f, ny = 0.f, nz = 0.f;
    if (best_t >= 0) {
        const float4 nn = reinterpret_cast<const float4*>(tris + 12 * (int64_t)best_t)[2];
        nx = nn.y;
        ny = nn.z;
        nz = nn.w;
    }
    const float dist = sqrtf(best) / h;
    const uint8_t sg = sign[v];
    const float scan_in = (sg == 1) ? 1.0f : 0.0f;
    const float t_off = jx * nx + jy * ny;
    const float s_out = 1.0f - 2.0f * scan_in;
    const float spread = spread_cnt ? 1.0f - sqrtf(sx * sx + sy * sy + sz * sz) / (float)spread_cnt : 0.0f;
    float* f = feats + 12 * i;
    f[0] = dist;
    f[1] = scan_in;
    f[2] = nx;
    f[3] = ny;
    f[4] = nz;
    f[6] = s_out * t_off - dist;  // f[5] and f[7]: axis_gather_kernel
    f[8] = log1pf((float)near_cnt);
    f[9] = spread;
    f[10] = risky ? 1.0f : 0.0f;
    f[11] = fminf(zhit / h, 2.0f);
    risk[i] = risky ? 1 : 0;
    wave_append(n_risky, risky);
}

// Triangles with their coordinates rotated so that axis `axis` (0 = x,
// 1 = y) becomes the third one: axis 0 maps (x, y, z) to (y, z, x), axis 1 to
// (z, x, y).  Both maps are cyclic (rotations), so a triangle's facing along
// the ray axis keeps its sign, and the column lists and the crossing test of
// the z scan apply unchanged to rays along x or y.  Vertices and normal
// alike.
__global__ void rotate_tris_kernel(const float* __restrict__ tris, int64_t nt, int axis, float* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const float4* src = reinterpret_cast<const float4*>(tris + 12 * t);
    const float4 a = src[0], b = src[1], c = src[2];
    const float v[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
    float r[12];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float X = v[3 * k], Y = v[3 * k + 1], Z = v[3 * k + 2];
        r[3 * k] = axis == 0 ? Y : Z;
        r[3 * k + 1] = axis == 0 ? Z : X;
        r[3 * k + 2] = axis == 0 ? X : Y;
    }
    float4* dst = reinterpret_cast<float4*>(out + 12 * t);
    dst[0] = make_float4(r[0], r[1], r[2], r[3]);
    dst[1] = make_float4(r[4], r[5], r[6], r[7]);
    dst[2] = make_float4(r[8], r[9], r[10], r[11]);
}

// Row lists for features 5 (axis 0, a ray along x) and 7 (axis 1, along y),
// over the rotated triangles and lattice (rotated_grid): row (iu, iv) holds
// exactly the triangles whose (u, v) box contains the row's jittered ray
// point ((mnu + iu + jx) h, (mnv + iv + jy) h), the box reject of the
// crossing test, with the same float predicate (the first and last index
// from an estimate and exact steps; the point moves monotonically with the
// index).  A triangle smaller than a voxel lands in about one row.
__device__ __forceinline__ float ray_at(int mn, int i, float j, float h) { return ((float)(mn + i) + j) * h; }

__device__ __forceinline__ int to_int_clamped(float v) { return (int)fminf(fmaxf(v, -2.0e9f), 2.0e9f); }

__device__ __forceinline__ void ray_rows(float lo, float hi, int mn, int d, float j, float h, int* first, int* last) {
    int i0 = to_int_clamped(ceilf(lo / h - j)) - mn;
    i0 = min(max(i0, 0), d);
    while (i0 > 0 && ray_at(mn, i0 - 1, j, h) >= lo) --i0;
    while (i0 < d && ray_at(mn, i0, j, h) < lo) ++i0;
    int i1 = to_int_clamped(floorf(hi / h - j)) - mn;
    i1 = min(max(i1, -1), d - 1);
    while (i1 < d - 1 && ray_at(mn, i1 + 1, j, h) <= hi) ++i1;
    while (i1 >= 0 && ray_at(mn, i1, j, h) > hi) --i1;
    *first = i0;
    *last = i1;
}

// A triangle covering at most kSmallRows rows is handled by one lane; a
// larger one goes to the `big` list, whose triangles grid-stride waves take
// (the lanes striding over the rows), so a few large triangles covering
// thousands of rows do not serialise on one lane, and many small ones do not
// leave 63 lanes of a wave idle.
constexpr int kSmallRows = 16;

__global__ void axis_rows_count_kernel(const float* __restrict__ rtris, int64_t nt, int mnu, int mnv, int du, int dv,
                                       float h, float jx, float jy, int4* __restrict__ cover,
                                       uint32_t* __restrict__ counts, int32_t* __restrict__ big,
                                       unsigned long long* n_big, unsigned long long* total) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long mine = 0;
    bool is_big = false;
    if (t < nt) {
        const Tri9 tr = load_tri(rtris, t);
        const float u0 = fminf(tr.ax, fminf(tr.bx, tr.cx)), u1 = fmaxf(tr.ax, fmaxf(tr.bx, tr.cx));
        const float v0 = fminf(tr.ay, fminf(tr.by, tr.cy)), v1 = fmaxf(tr.ay, fmaxf(tr.by, tr.cy));
        int4 cv;
        ray_rows(u0, u1, mnu, du, jx, h, &cv.x, &cv.y);
        ray_rows(v0, v1, mnv, dv, jy, h, &cv.z, &cv.w);
        cover[t] = cv;
        if (cv.y >= cv.x && cv.w >= cv.z) {
            mine = (unsigned long long)(cv.y - cv.x + 1) * (unsigned long long)(cv.w - cv.z + 1);
            is_big = mine > (unsigned long long)kSmallRows;
            if (!is_big)
                for (int iv = cv.z; iv <= cv.w; ++iv)
                    for (int iu = cv.x; iu <= cv.y; ++iu) atomicAdd(counts + (int64_t)iv * du + iu, 1u);
        }
    }
    const long long slot = wave_append(n_big, is_big);
    if (is_big) big[slot] = (int32_t)t;
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
    if (__lane_id() == 0 && mine) atomicAdd(total, mine);
}

// The big list's rows: one wave per triangle, the waves striding over the list.
__global__ __launch_bounds__(256) void axis_rows_big_count_kernel(const int4* __restrict__ cover,
                                                                  const int32_t* __restrict__ big,
                                                                  const unsigned long long* n_big, int du,
                                                                  uint32_t* __restrict__ counts) {
    const int64_t nb = (int64_t)*n_big;
    for (int64_t q = (int64_t)blockIdx.x * 4 + threadIdx.y; q < nb; q += (int64_t)gridDim.x * 4) {
        const int4 cv = cover[big[q]];
        const int w = cv.y - cv.x + 1;
        const int64_t area = (int64_t)w * (cv.w - cv.z + 1);
        for (int64_t k = threadIdx.x; k < area; k += 64)
            atomicAdd(counts + (int64_t)(cv.z + k / w) * du + cv.x + k % w, 1u);
    }
}

__global__ void axis_rows_fill_kernel(int64_t nt, int du, const int4* __restrict__ cover,
                                      const int64_t* __restrict__ start, uint32_t* __restrict__ cursor,
                                      int32_t* __restrict__ ids) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int4 cv = cover[t];
    if (cv.y < cv.x || cv.w < cv.z) return;
    if ((int64_t)(cv.y - cv.x + 1) * (cv.w - cv.z + 1) > kSmallRows) return;  // the big list's
    for (int iv = cv.z; iv <= cv.w; ++iv)
        for (int iu = cv.x; iu <= cv.y; ++iu) {
            const int64_t r = (int64_t)iv * du + iu;
            ids[start[r] + atomicAdd(cursor + r, 1u)] = (int32_t)t;
        }
}

__global__ __launch_bounds__(256) void axis_rows_big_fill_kernel(const int4* __restrict__ cover,
                                                                 const int32_t* __restrict__ big,
                                                                 const unsigned long long* n_big, int du,
                                                                 const int64_t* __restrict__ start,
                                                                 uint32_t* __restrict__ cursor,
                                                                 int32_t* __restrict__ ids) {
    const int64_t nb = (int64_t)*n_big;
    for (int64_t q = (int64_t)blockIdx.x * 4 + threadIdx.y; q < nb; q += (int64_t)gridDim.x * 4) {
        const int32_t t = big[q];
        const int4 cv = cover[t];
        const int w = cv.y - cv.x + 1;
        const int64_t area = (int64_t)w * (cv.w - cv.z + 1);
        for (int64_t k = threadIdx.x; k < area; k += 64) {
            const int64_t r = (int64_t)(cv.z + k / w) * du + cv.x + k % w;
            ids[start[r] + atomicAdd(cursor + r, 1u)] = t;
        }
    }
}

// Features 5 / 7: the signed crossing count of the candidate's row ray, with
// the z scan's crossing test and rule (a crossing counts as below the voxel
// when its height is < the voxel centre's).  1 when the crossings below the
// voxel do not cancel (inside by that ray), 0 when they do, 0.5 when the
// crossings of the whole line do not cancel (an open line gives no answer).
// With the z scan sign (feature 1) these are the three axis parity counts.
// The sum does not depend on the list order.
__global__ void axis_parity_kernel(const float* __restrict__ rtris, const int64_t* __restrict__ start,
                                   const int32_t* __restrict__ ids, const int64_t* __restrict__ list, int64_t n,
                                   int dx, int dy, int axis, int mnu, int mnv, int mnw, int du, float h, float jx,
                                   float jy, float* __restrict__ feats, int slot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t v = list[i];
    const int x = (int)(v % dx);
    const int64_t r = v / dx;
    const int y = (int)(r % dy);
    const int z = (int)(r / dy);
    const int iu = axis == 0 ? y : z, iv = axis == 0 ? z : x, iw = axis == 0 ? x : y;
    const float ru = ray_at(mnu, iu, jx, h), rv = ray_at(mnv, iv, jy, h);
    const float pw = lattice(mnw, iw, h);
    const int64_t row = (int64_t)iv * du + iu;
    const int64_t e0 = start[row], e1 = start[row + 1];
    int total = 0, below = 0;
    for (int64_t eb = e0; eb < e1; eb += kFeatBatch) {
        // every index is clamped into [e0, e1), so the loads need no guard
        int tid[kFeatBatch];
#pragma unroll
        for (int j = 0; j < kFeatBatch; ++j) tid[j] = ids[min(eb + j, e1 - 1)];
#pragma unroll
        for (int j = 0; j < kFeatBatch; ++j) {
            if (eb + j >= e1) continue;
            const Tri9 t = load_tri(rtris, tid[j]);
            const float d0 = (t.bx - t.ax) * (rv - t.ay) - (t.by - t.ay) * (ru - t.ax);
            const float d1 = (t.cx - t.bx) * (rv - t.by) - (t.cy - t.by) * (ru - t.bx);
            const float d2 = (t.ax - t.cx) * (rv - t.cy) - (t.ay - t.cy) * (ru - t.cx);
            const bool pos = (d0 > 0.0f && d1 > 0.0f && d2 > 0.0f);
            const bool neg = (d0 < 0.0f && d1 < 0.0f && d2 < 0.0f);
            if (!pos && !neg) continue;
            const float area2 = d0 + d1 + d2;
            if (fabsf(area2) < 1e-12f) continue;
            const float wh = (d1 * t.az + d2 * t.bz + d0 * t.cz) / area2;
            const int sgn = pos ? 1 : -1;
            total += sgn;
            if (wh < pw) below += sgn;
        }
    }
    feats[12 * i + slot] = (total != 0) ? 0.5f : ((below != 0) ? 1.0f : 0.0f);
}

__global__ void fill_feature_kernel(float* __restrict__ feats, int64_t n, int slot, float value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) feats[12 * i + slot] = value;
}

// Risky candidates the ranked probes did not cover (the certificate's
// fallback list).
__global__ void unprobed_risky_kernel(const int64_t* __restrict__ list, const uint8_t* __restrict__ risk, int64_t n,
                                      const unsigned* __restrict__ bits, int64_t* __restrict__ out, int64_t cap,
                                      unsigned long long* counter) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t iters = (n + stride - 1) / stride;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t i = tid + it * stride;
        bool want = false;
        int64_t v = 0;
        if (i < n && risk[i]) {
            v = list[i];
            want = ((bits[v >> 5] >> (v & 31)) & 1u) == 0;
        }
        long long slot = wave_append(counter, want);
        if (want && slot < cap) out[slot] = v;
    }
}

// score = planar margin (feature 6): the rule the network starts from
__global__ void margin_score_kernel(const float* __restrict__ feats, int64_t n, float* __restrict__ score) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) score[i] = feats[12 * i + 6];
}

// Rankings without a model: a uniform position in [0, 0.5) drawn from a
// splitmix64 hash of the candidate's voxel index (not of its list position,
// which the atomic compaction leaves run-dependent), plus the certificate flag
// (feature 10) when `feats` is given: flagged candidates then score in
// [1, 1.5) and rank first, each class in hashed order.  Without `feats` the
// top k are a blind uniform subset of the candidates.
__global__ void hash_score_kernel(const float* __restrict__ feats, const int64_t* __restrict__ cand, int64_t n,
                                  float* __restrict__ score) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = 0x9E3779B97F4A7C15ull * (unsigned long long)(cand[i] + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float u = (float)(z >> 40) * 0x1.0p-24f;  // [0, 1)
    score[i] = (feats ? feats[12 * i + 10] : 0.0f) + 0.5f * u;
}

constexpr int kBins = 4096;

__device__ __forceinline__ int score_bin(float s, float lo, float inv_w) {
    float b = (s - lo) * inv_w;
    b = fminf(fmaxf(b, 0.0f), (float)(kBins - 1));
    return (int)b;
}

__global__ void hist_kernel(const float* __restrict__ score, int64_t n, float lo, float inv_w,
                            unsigned* __restrict__ hist) {
    __shared__ unsigned sh[kBins];
    for (int b = threadIdx.x; b < kBins; b += blockDim.x) sh[b] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&sh[score_bin(score[i], lo, inv_w)], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kBins; b += blockDim.x)
        if (sh[b]) atomicAdd(&hist[b], sh[b]);
}

__global__ void select_kernel(const float* __restrict__ score, const int64_t* __restrict__ list, int64_t n,
                              float lo, float inv_w, int min_bin, int64_t* __restrict__ out, int64_t cap,
                              unsigned long long* counter) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t iters = (n + stride - 1) / stride;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t i = tid + it * stride;
        const bool want = (i < n) && score_bin(score[i], lo, inv_w) >= min_bin;
        long long slot = wave_append(counter, want);
        if (want && slot < cap) out[slot] = list[i];
    }
}

__global__ void mark_probed_kernel(const int64_t* __restrict__ list, int64_t n, unsigned* __restrict__ bits) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int64_t v = list[i];
        atomicOr(&bits[v >> 5], 1u << (v & 31));
    }
}

// Commit exact decisions; append voxels whose scanline sign was wrong.
__global__ void commit_kernel(const double* __restrict__ w, const int64_t* __restrict__ list, int64_t n,
                              uint8_t* __restrict__ sign, int64_t* __restrict__ mism, int64_t cap,
                              unsigned long long* counter) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t iters = (n + stride - 1) / stride;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t i = tid + it * stride;
        bool bad = false;
        int64_t v = 0;
        if (i < n) {
            v = list[i];
            const uint8_t exact = (fabs(w[i]) >= 0.5) ? 1 : 0;
            bad = sign[v] != exact;
            sign[v] = exact;
        }
        long long slot = wave_append(counter, bad);
        if (bad && slot < cap) mism[slot] = v;
    }
}

// Unprobed 26-neighbours of every mismatch, each claimed once via the bitmap.
__global__ void grow_kernel(const int64_t* __restrict__ mism, int64_t m, int dx, int dy, int dz,
                            unsigned* __restrict__ bits, int64_t* __restrict__ out, int64_t cap,
                            unsigned long long* counter) {
    const int64_t total = m * 27;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t iters = (total + stride - 1) / stride;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t k = tid + it * stride;
        bool want = false;
        int64_t u = 0;
        if (k < total) {
            const int64_t v = mism[k / 27];
            const int o = (int)(k % 27);
            const int ox = o % 3 - 1, oy = (o / 3) % 3 - 1, oz = o / 9 - 1;
            const int x = (int)(v % dx) + ox;
            const int64_t r = v / dx;
            const int y = (int)(r % dy) + oy;
            const int z = (int)(r / dy) + oz;
            if (o != 13 && x >= 0 && x < dx && y >= 0 && y < dy && z >= 0 && z < dz) {
                u = ((int64_t)z * dy + y) * dx + x;
                const unsigned bit = 1u << (u & 31);
                want = (atomicOr(&bits[u >> 5], bit) & bit) == 0;
            }
        }
        long long slot = wave_append(counter, want);
        if (want && slot < cap) out[slot] = u;
    }
}

}  // namespace

hipError_t launch_candidates(const float* d, int64_t n, float thr, int64_t* list, int64_t cap,
                             unsigned long long* counter, hipStream_t s) {
    const int64_t chunks = (n / 4 + (int64_t)kCandBlock * kCandPer - 1) / ((int64_t)kCandBlock * kCandPer);
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(chunks, 8192));
    candidates_kernel<<<blocks, kCandBlock, 0, s>>>(d, n, thr, list, cap, counter);
    return hipGetLastError();
}

size_t zsort_scratch_bytes(int64_t entries, int64_t nsup, const int64_t* start) {
    size_t bytes = 0;
    if (rocprim::segmented_radix_sort_keys(nullptr, bytes, (const unsigned long long*)nullptr,
                                           (unsigned long long*)nullptr, (unsigned)entries, (unsigned)nsup, start,
                                           start + 1) != hipSuccess)
        return 0;
    return bytes + 256;
}

hipError_t launch_zsort_lists(const float* tris, int64_t nt, const CsrBuffers& b, int64_t entries, int64_t nsup,
                              float2* zr, unsigned* max_ext, unsigned long long* keys_tmp, unsigned long long* keys,
                              void* scratch, size_t scratch_bytes, hipStream_t s) {
    if (nt <= 0 || entries <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(max_ext, 0, sizeof(unsigned), s);
    if (e != hipSuccess) return e;
    tri_zrange_kernel<<<(unsigned)std::min<int64_t>((nt + 255) / 256, 2048), 256, 0, s>>>(tris, nt, zr, max_ext);
    zsort_keys_kernel<<<(unsigned)((entries + 255) / 256), 256, 0, s>>>(b.ids, entries, zr, keys_tmp);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = scratch_bytes;
    return rocprim::segmented_radix_sort_keys(scratch, bytes, keys_tmp, keys, (unsigned)entries, (unsigned)nsup,
                                              b.start, b.start + 1, 0, 64, s);
}

hipError_t launch_features(const float* tris, int64_t nt, const CsrBuffers& b, const float2* zr,
                           const unsigned long long* keys, const unsigned* max_ext, const uint8_t* sign,
                           const int64_t* list, int64_t n, const Grid& g, int sup, float jx, float jy, float* feats,
                           uint8_t* risk, unsigned long long* n_risky, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int sdx = (g.dims[0] + sup - 1) / sup;
    features_kernel<<<(unsigned)((n + 127) / 128), 128, 0, s>>>(tris, b.xyab, zr, b.start, keys, max_ext, sign, list, n,
                                                              g.mn[0], g.mn[1], g.mn[2], g.dims[0], g.dims[1], g.h, sup,
                                                              sdx, jx, jy, feats, risk, n_risky);
    return hipGetLastError();
}

hipError_t launch_unprobed_risky(const int64_t* list, const uint8_t* risk, int64_t n, const unsigned* bits,
                                 int64_t* out, int64_t cap, unsigned long long* counter, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    unprobed_risky_kernel<<<grid_for(n), kBlock, 0, s>>>(list, risk, n, bits, out, cap, counter);
    return hipGetLastError();
}

hipError_t launch_rotate_tris(const float* tris, int64_t nt, int axis, float* out, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    rotate_tris_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(tris, nt, axis, out);
    return hipGetLastError();
}

hipError_t launch_axis_rows_count(const float* rtris, int64_t nt, const Grid& r, float jx, float jy, int4* cover,
                                  uint32_t* counts, int32_t* big, unsigned long long* n_big,
                                  unsigned long long* total, hipStream_t s) {
    hipError_t e = hipMemsetAsync(counts, 0, sizeof(uint32_t) * (size_t)r.dims[0] * r.dims[1], s);
    if (e == hipSuccess) e = hipMemsetAsync(n_big, 0, sizeof(unsigned long long), s);
    if (e != hipSuccess || nt <= 0) return e;
    axis_rows_count_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(rtris, nt, r.mn[0], r.mn[1], r.dims[0],
                                                                        r.dims[1], r.h, jx, jy, cover, counts, big,
                                                                        n_big, total);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    axis_rows_big_count_kernel<<<1024, dim3(64, 4), 0, s>>>(cover, big, n_big, r.dims[0], counts);
    return hipGetLastError();
}

hipError_t launch_axis_rows_fill(int64_t nt, const Grid& r, const int4* cover, const int32_t* big,
                                 const unsigned long long* n_big, const int64_t* start, uint32_t* cursor, int32_t* ids,
                                 hipStream_t s) {
    hipError_t e = hipMemsetAsync(cursor, 0, sizeof(uint32_t) * (size_t)r.dims[0] * r.dims[1], s);
    if (e != hipSuccess || nt <= 0) return e;
    axis_rows_fill_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(nt, r.dims[0], cover, start, cursor, ids);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    axis_rows_big_fill_kernel<<<1024, dim3(64, 4), 0, s>>>(cover, big, n_big, r.dims[0], start, cursor, ids);
    return hipGetLastError();
}

hipError_t launch_axis_parity(const float* rtris, const int64_t* start, const int32_t* ids, const int64_t* list,
                              int64_t n, const Grid& g, int axis, float jx, float jy, float* feats, int slot,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const Grid r = rotated_grid(g, axis);
    axis_parity_kernel<<<(unsigned)((n + 127) / 128), 128, 0, s>>>(rtris, start, ids, list, n, g.dims[0], g.dims[1],
                                                                   axis, r.mn[0], r.mn[1], r.mn[2], r.dims[0], g.h,
                                                                   jx, jy, feats, slot);
    return hipGetLastError();
}

hipError_t launch_fill_feature(float* feats, int64_t n, int slot, float value, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    fill_feature_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(feats, n, slot, value);
    return hipGetLastError();
}

hipError_t launch_margin_score(const float* feats, int64_t n, float* score, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    margin_score_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(feats, n, score);
    return hipGetLastError();
}

hipError_t launch_hash_score(const float* feats, const int64_t* cand, int64_t n, float* score, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hash_score_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(feats, cand, n, score);
    return hipGetLastError();
}

int score_bins() { return kBins; }

hipError_t launch_score_hist(const float* score, int64_t n, float lo, float hi, unsigned* hist, hipStream_t s) {
    hipError_t e = hipMemsetAsync(hist, 0, sizeof(unsigned) * kBins, s);
    if (e != hipSuccess || n <= 0) return e;
    hist_kernel<<<grid_for(n, 1024), 1024, 0, s>>>(score, n, lo, (float)kBins / (hi - lo), hist);
    return hipGetLastError();
}

hipError_t launch_select(const float* score, const int64_t* list, int64_t n, float lo, float hi, int min_bin,
                         int64_t* out, int64_t cap, unsigned long long* counter, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    select_kernel<<<grid_for(n), kBlock, 0, s>>>(score, list, n, lo, (float)kBins / (hi - lo), min_bin, out, cap,
                                                counter);
    return hipGetLastError();
}

hipError_t launch_mark_probed(const int64_t* list, int64_t n, unsigned* bits, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    mark_probed_kernel<<<grid_for(n), kBlock, 0, s>>>(list, n, bits);
    return hipGetLastError();
}

hipError_t launch_commit(const double* w, const int64_t* list, int64_t n, uint8_t* sign, int64_t* mism, int64_t cap,
                         unsigned long long* counter, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    commit_kernel<<<grid_for(n), kBlock, 0, s>>>(w, list, n, sign, mism, cap, counter);
    return hipGetLastError();
}

hipError_t launch_grow(const int64_t* mism, int64_t m, const int32_t dims[3], unsigned* bits, int64_t* out,
                       int64_t cap, unsigned long long* counter, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    grow_kernel<<<grid_for(m * 27), kBlock, 0, s>>>(mism, m, dims[0], dims[1], dims[2], bits, out, cap, counter);
    return hipGetLastError();
}

}  // namespace rhz
