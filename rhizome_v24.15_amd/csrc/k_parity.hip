// k_parity.hip -- [B] scanline parity signs.
//
// Candidate lists: the reference builds a CSR of triangles per 4x4-column
// supercell on the host (cuda_rasterizer.cu:824-882).  Here the same coverage
// rule (XY AABB, +-1 column slack, truncating integer division) is evaluated
// on the device: a counting pass, an exclusive scan, and a fill pass.  The
// order of ids inside a supercell is arbitrary; the parity result does not
// depend on it (hits are sorted, overflow is a count).
//
// Parity: one lane per (x, y) column casts the jittered z-ray
// (cuda_rasterizer.cu:279-368).  Instead of sorting FP32 hit heights and
// comparing them with every voxel centre, each hit is converted once to the
// first voxel index k whose centre lies strictly above it (the exact
// predicate hit < (min_z + z) * h of :361), so the walk down the column only
// compares integers.  k values live in LDS in a lane-interleaved layout
// (bank-conflict free), are insertion-sorted, and the column is written as
// one byte per voxel.
//
// The walk is latency-bound (candidate id -> XY box -> triangle are three
// dependent loads), so the ids and boxes of kWalkBatch entries are fetched
// together before any of them is tested; entries are still tested in list
// order, so the overflow cut-off sees exactly the same sequence.  Keys are
// 16-bit when they fit (dz <= 32767), which halves the LDS per block and
// doubles the waves a CU keeps in flight.
#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

constexpr int kBlock = 256;
constexpr int kParityBlock = 64;     // one wave: 16 or 32 KB of hit keys in LDS
constexpr int kParityMaxHits = 128;  // LDS capacity per lane (reference limit)
constexpr int kWalkBatch = 8;        // candidate entries whose loads are issued together
constexpr int kShortKeyMaxDz = 32767;  // (dz << 1) | 1 fits 16 bits

__device__ __forceinline__ int floor_div_trunc(int a, int b) { return a / b; }  // C semantics

__device__ __forceinline__ void tri_cover(const Tri9& t, float h, int sup, int mnx, int mny, int sdx,
                                          int sdy, float4& ab, int4& cov) {
    float minx = fminf(t.ax, fminf(t.bx, t.cx)), maxx = fmaxf(t.ax, fmaxf(t.bx, t.cx));
    float miny = fminf(t.ay, fminf(t.by, t.cy)), maxy = fmaxf(t.ay, fmaxf(t.by, t.cy));
    ab = make_float4(minx, miny, maxx, maxy);
    auto to_int = [](float v) { return (int)fminf(fmaxf(v, -2.0e9f), 2.0e9f); };
    int cx0 = to_int(floorf(minx / h)) - mnx - 1;
    int cx1 = to_int(ceilf(maxx / h)) - mnx + 1;
    int cy0 = to_int(floorf(miny / h)) - mny - 1;
    int cy1 = to_int(ceilf(maxy / h)) - mny + 1;
    int sx0 = max(0, floor_div_trunc(cx0, sup));
    int sx1 = min(sdx - 1, floor_div_trunc(cx1, sup));
    int sy0 = max(0, floor_div_trunc(cy0, sup));
    int sy1 = min(sdy - 1, floor_div_trunc(cy1, sup));
    cov = make_int4(sx0, sx1, sy0, sy1);
}

__global__ void csr_count_kernel(const float* __restrict__ tris, int64_t nt, float h, int sup, int mnx,
                                 int mny, int sdx, int sdy, float4* __restrict__ xyab,
                                 int4* __restrict__ cover, uint32_t* __restrict__ counts,
                                 unsigned long long* total) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long mine = 0;
    if (t < nt) {
        Tri9 tr = load_tri(tris, t);
        float4 ab;
        int4 cv;
        tri_cover(tr, h, sup, mnx, mny, sdx, sdy, ab, cv);
        xyab[t] = ab;
        cover[t] = cv;
        if (cv.y >= cv.x && cv.w >= cv.z) {
            for (int sy = cv.z; sy <= cv.w; ++sy)
                for (int sx = cv.x; sx <= cv.y; ++sx) atomicAdd(counts + (int64_t)sy * sdx + sx, 1u);
            mine = (unsigned long long)(cv.y - cv.x + 1) * (unsigned long long)(cv.w - cv.z + 1);
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
    if (__lane_id() == 0 && mine) atomicAdd(total, mine);
}

__global__ void csr_fill_kernel(int64_t nt, int sdx, const int4* __restrict__ cover,
                                const int64_t* __restrict__ start, uint32_t* __restrict__ cursor,
                                int32_t* __restrict__ ids) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int4 cv = cover[t];
    if (cv.y < cv.x || cv.w < cv.z) return;
    for (int sy = cv.z; sy <= cv.w; ++sy)
        for (int sx = cv.x; sx <= cv.y; ++sx) {
            int64_t s = (int64_t)sy * sdx + sx;
            uint32_t slot = atomicAdd(cursor + s, 1u);
            ids[start[s] + slot] = (int32_t)t;
        }
}

// First voxel index z in [0, dz] with hit < (mnz + z) * h (dz if none).
__device__ __forceinline__ int toggle_index(float hit, int mnz, int dz, float h) {
    float e = floorf(hit / h);
    e = fminf(fmaxf(e, -2.0e9f), 2.0e9f);
    int k = (int)e - mnz + 1;
    k = min(max(k, 0), dz);
    while (k > 0 && hit < lattice(mnz, k - 1, h)) --k;
    while (k < dz && !(hit < lattice(mnz, k, h))) ++k;
    return k;
}

// SIGNED = false: reference parity (inside = odd number of hits below).
// SIGNED = true: signed crossing count along the same jittered ray, i.e. the
// winding number of the ray point w = sum over hits above of +1 (triangle
// facing +z: the ray leaves the solid) / -1 (facing -z).  Inside = w != 0,
// which also classifies overlapping shells (w = 2) that parity calls outside.
// A column whose crossings do not cancel (w above the top != 0) is deferred
// like an odd one.  Key = uint16_t when dz <= kShortKeyMaxDz, else int.
template <bool SIGNED, typename Key>
__global__ __launch_bounds__(kParityBlock) void parity_kernel(
    const float* __restrict__ tris, const float4* __restrict__ xyab, const int64_t* __restrict__ start,
    const int32_t* __restrict__ ids, int mnx, int mny, int mnz, int dx, int dy, int dz, float h, int sup,
    int sdx, float jx, float jy, int max_hits, uint8_t* __restrict__ sign, int* __restrict__ flag) {
    __shared__ Key kbuf[kParityMaxHits][kParityBlock];
    const int lane = threadIdx.x;
    const int64_t col = (int64_t)blockIdx.x * kParityBlock + lane;
    const int64_t ncols = (int64_t)dx * dy;
    if (col >= ncols) return;
    const int cx = (int)(col % dx), cy = (int)(col / dx);
    const float rx = ((float)(mnx + cx) + jx) * h;
    const float ry = ((float)(mny + cy) + jy) * h;

    int n = 0, total = 0;
    bool bad = false;
    const int64_t s = (int64_t)(cy / sup) * sdx + (cx / sup);
    const int64_t e0 = start[s], e1 = start[s + 1];
    for (int64_t eb = e0; eb < e1 && !bad; eb += kWalkBatch) {
        // every index is clamped into [e0, e1), so the loads need no guard
        int tid[kWalkBatch];
#pragma unroll
        for (int j = 0; j < kWalkBatch; ++j) tid[j] = ids[min(eb + j, e1 - 1)];
        bool in_box[kWalkBatch];
#pragma unroll
        for (int j = 0; j < kWalkBatch; ++j) {
            const float4 ab = xyab[tid[j]];
            in_box[j] = (eb + j < e1) && !(rx < ab.x || rx > ab.z || ry < ab.y || ry > ab.w);
        }
#pragma unroll
        for (int j = 0; j < kWalkBatch; ++j) {
            if (!in_box[j] || bad) continue;
            Tri9 t = load_tri(tris, tid[j]);
            float d0 = (t.bx - t.ax) * (ry - t.ay) - (t.by - t.ay) * (rx - t.ax);
            float d1 = (t.cx - t.bx) * (ry - t.by) - (t.cy - t.by) * (rx - t.bx);
            float d2 = (t.ax - t.cx) * (ry - t.cy) - (t.ay - t.cy) * (rx - t.cx);
            bool pos = (d0 > 0.0f && d1 > 0.0f && d2 > 0.0f);
            bool neg = (d0 < 0.0f && d1 < 0.0f && d2 < 0.0f);
            if (!pos && !neg) continue;
            float area2 = d0 + d1 + d2;
            if (fabsf(area2) < 1e-12f) continue;
            float zh = (d1 * t.az + d2 * t.bz + d0 * t.cz) / area2;
            if (n < max_hits) {
                const int k = toggle_index(zh, mnz, dz, h);
                kbuf[n][lane] = (Key)(SIGNED ? ((k << 1) | (pos ? 1 : 0)) : k);
                total += pos ? 1 : -1;
                ++n;
            } else {
                bad = true;  // the later entries of this batch are skipped, as the list walk ends here
            }
        }
    }

    const int64_t zs = (int64_t)dx * dy;
    uint8_t* out = sign + col;
    if (bad || (n & 1) || (SIGNED && total != 0)) {
        for (int z = 0; z < dz; ++z) out[(int64_t)z * zs] = 2;
        *flag = 1;
        return;
    }
    for (int i = 1; i < n; ++i) {
        int key = kbuf[i][lane];
        int j = i - 1;
        while (j >= 0 && kbuf[j][lane] > key) {
            kbuf[j + 1][lane] = kbuf[j][lane];
            --j;
        }
        kbuf[j + 1][lane] = (Key)key;
    }
    const int shift = SIGNED ? 1 : 0;
    int p = 0;
    int next = (n > 0) ? ((int)kbuf[0][lane] >> shift) : dz;
    uint8_t inside = 0;
    int below = 0;  // signed crossings below the voxel centre
    for (int z = 0; z < dz; ++z) {
        while (next <= z) {
            if (SIGNED)
                below += (kbuf[p][lane] & 1) ? 1 : -1;
            else
                inside ^= 1;
            ++p;
            next = (p < n) ? ((int)kbuf[p][lane] >> shift) : dz;
        }
        out[(int64_t)z * zs] = SIGNED ? (uint8_t)(below != 0) : inside;
    }
}

template <bool SIGNED, typename Key>
void launch_parity_t(unsigned blocks, const float* tris, const Grid& g, int sup, int sdx, const CsrBuffers& b,
                     float jx, float jy, int max_hits, uint8_t* sign, int* flag, hipStream_t s) {
    parity_kernel<SIGNED, Key><<<blocks, kParityBlock, 0, s>>>(tris, b.xyab, b.start, b.ids, g.mn[0], g.mn[1],
                                                               g.mn[2], g.dims[0], g.dims[1], g.dims[2], g.h, sup,
                                                               sdx, jx, jy, max_hits, sign, flag);
}

}  // namespace

int64_t num_supercells(const Grid& g, int sup) {
    int64_t sdx = (g.dims[0] + sup - 1) / sup, sdy = (g.dims[1] + sup - 1) / sup;
    return sdx * sdy;
}

hipError_t launch_csr_count(const float* tris, int64_t nt, const Grid& g, int sup, const CsrBuffers& b,
                            unsigned long long* total, hipStream_t s) {
    int sdx = (g.dims[0] + sup - 1) / sup, sdy = (g.dims[1] + sup - 1) / sup;
    hipError_t e = hipMemsetAsync(b.counts, 0, sizeof(uint32_t) * (size_t)sdx * sdy, s);
    if (e != hipSuccess) return e;
    if (nt <= 0) return hipSuccess;
    csr_count_kernel<<<(unsigned)((nt + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        tris, nt, g.h, sup, g.mn[0], g.mn[1], sdx, sdy, b.xyab, b.cover, b.counts, total);
    return hipGetLastError();
}

hipError_t launch_csr_fill(int64_t nt, const Grid& g, int sup, const CsrBuffers& b, hipStream_t s) {
    int sdx = (g.dims[0] + sup - 1) / sup, sdy = (g.dims[1] + sup - 1) / sup;
    hipError_t e = hipMemsetAsync(b.cursor, 0, sizeof(uint32_t) * (size_t)sdx * sdy, s);
    if (e != hipSuccess) return e;
    if (nt <= 0) return hipSuccess;
    csr_fill_kernel<<<(unsigned)((nt + kBlock - 1) / kBlock), kBlock, 0, s>>>(nt, sdx, b.cover, b.start,
                                                                             b.cursor, b.ids);
    return hipGetLastError();
}

hipError_t launch_parity(const float* tris, const Grid& g, int sup, const CsrBuffers& b, float jx, float jy,
                         int max_hits, uint8_t* sign, int* flag, hipStream_t s, bool signed_count) {
    if (max_hits > kParityMaxHits) return hipErrorInvalidValue;
    int sdx = (g.dims[0] + sup - 1) / sup;
    int64_t ncols = (int64_t)g.dims[0] * g.dims[1];
    const unsigned blocks = (unsigned)((ncols + kParityBlock - 1) / kParityBlock);
    const bool short_keys = g.dims[2] <= kShortKeyMaxDz;
    if (signed_count) {
        if (short_keys)
            launch_parity_t<true, uint16_t>(blocks, tris, g, sup, sdx, b, jx, jy, max_hits, sign, flag, s);
        else
            launch_parity_t<true, int>(blocks, tris, g, sup, sdx, b, jx, jy, max_hits, sign, flag, s);
    } else {
        if (short_keys)
            launch_parity_t<false, uint16_t>(blocks, tris, g, sup, sdx, b, jx, jy, max_hits, sign, flag, s);
        else
            launch_parity_t<false, int>(blocks, tris, g, sup, sdx, b, jx, jy, max_hits, sign, flag, s);
    }
    return hipGetLastError();
}

}  // namespace rhz
