// k_grid.hip -- streaming passes over the voxel grid (HBM-bound byte work):
// sentinel fill, sign apply, far-sentinel audit and clamp, unknown-voxel
// compaction, patch-box marking, probe positions, and an exclusive scan.
#include <algorithm>

#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

constexpr int kBlock = 256;

inline unsigned stream_blocks(int64_t work_items, int per_block = kBlock) {
    int64_t b = (work_items + per_block - 1) / per_block;
    // 256 CUs x 8 resident blocks is enough to saturate HBM for these passes.
    if (b > 2048 * 4) b = 2048 * 4;
    if (b < 1) b = 1;
    return (unsigned)b;
}

// [init] d[i] = v (reference gpu_fill_float_kernel, cuda_rasterizer.cu:194-203),
// 16-byte stores.
__global__ void zero_i32_kernel(int* __restrict__ p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0;
}

__global__ void fill_kernel(float* __restrict__ d, int64_t n, float v) {
    typedef float vf4 __attribute__((ext_vector_type(4)));
    const int64_t n4 = n >> 2;
    vf4 v4 = {v, v, v, v};
    vf4* d4 = reinterpret_cast<vf4*>(d);
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        __builtin_nontemporal_store(v4, d4 + i);
    int64_t t = (n4 << 2) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = v;
}

// [F] signed = inside ? -d : d, with d > 1e29 -> 1e6 (cuda_rasterizer.cu:650-661).
__device__ __forceinline__ float apply_one(float d, uint8_t s) { return apply_sign(d, s); }

__global__ void apply_kernel(float* __restrict__ d, const uint8_t* __restrict__ sign, int64_t n) {
    const int64_t n4 = n >> 2;
    float4* d4 = reinterpret_cast<float4*>(d);
    const uchar4* s4 = reinterpret_cast<const uchar4*>(sign);
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        float4 v = d4[i];
        uchar4 s = s4[i];
        v.x = apply_one(v.x, s.x);
        v.y = apply_one(v.y, s.y);
        v.z = apply_one(v.z, s.z);
        v.w = apply_one(v.w, s.w);
        d4[i] = v;
    }
    int64_t t = (n4 << 2) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = apply_one(d[t], sign[t]);
}

// [D3] safe clamp (cuda_rasterizer.cu:544-556).  intended = 1 stores the
// magnitude so [F] ships -bw inside; intended = 0 stores the reference's
// pre-signed value.
__global__ void clamp_far_kernel(float* __restrict__ d, const uint8_t* __restrict__ sign, int64_t n,
                                 float bw, int intended) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (d[i] >= kFarTest) d[i] = intended ? bw : ((sign[i] == 1) ? -bw : bw);
    }
}

// Row r of a field with rpp rows per plane sits at plane (r / rpp) and row
// (r % rpp) of that plane.
__device__ __forceinline__ int64_t row_offset(int64_t r, int rpp, int64_t pitch, int64_t plane) {
    const int64_t z = r / rpp;
    return z * plane + (r - z * rpp) * pitch;
}

// Rows of `len` floats copied between two layouts (a block per group of
// rows, threads along the row: both sides coalesced).
__global__ void copy_rows_kernel(const float* __restrict__ src, int64_t sp, int64_t splane, float* __restrict__ dst,
                                 int64_t dp, int64_t dplane, int len, int rpp, int64_t rows) {
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const float* a = src + row_offset(r, rpp, sp, splane);
        float* b = dst + row_offset(r, rpp, dp, dplane);
        for (int x = threadIdx.x; x < len; x += blockDim.x) b[x] = a[x];
    }
}

// copy_rows_kernel into the dense field (row r at r * len) with apply_one on
// the way, the sign read at the same dense index.
__global__ void copy_apply_rows_kernel(const float* __restrict__ src, int64_t sp, int64_t splane,
                                       float* __restrict__ dst, const uint8_t* __restrict__ sign, int len, int rpp,
                                       int64_t rows) {
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const float* a = src + row_offset(r, rpp, sp, splane);
        float* b = dst + r * len;
        const uint8_t* g = sign + r * len;
        for (int x = threadIdx.x; x < len; x += blockDim.x) b[x] = apply_one(a[x], g[x]);
    }
}

// Sentinel count over rows of `len` floats (the padding between rows and
// between planes is not counted).
__global__ void count_far_rows_kernel(const float* __restrict__ d, int64_t p, int64_t plane, int len, int rpp,
                                      int64_t rows, unsigned long long* counter) {
    unsigned c = 0;
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const float* a = d + row_offset(r, rpp, p, plane);
        for (int x = threadIdx.x; x < len; x += blockDim.x) c += (a[x] >= kFarTest);
    }
    for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (__lane_id() == 0 && c) atomicAdd(counter, (unsigned long long)c);
}

// [D2] audit: count d >= 1e29 (cuda_rasterizer.cu:523-534); one atomic per wave.
__global__ void count_far_kernel(const float* __restrict__ d, int64_t n, unsigned long long* counter) {
    const int64_t n4 = n >> 2;
    const float4* d4 = reinterpret_cast<const float4*>(d);
    unsigned c = 0;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        float4 v = d4[i];
        c += (v.x >= kFarTest) + (v.y >= kFarTest) + (v.z >= kFarTest) + (v.w >= kFarTest);
    }
    int64_t t = (n4 << 2) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) c += (d[t] >= kFarTest);
    for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (__lane_id() == 0 && c) atomicAdd(counter, (unsigned long long)c);
}

// Compaction of unknown voxels (sign == 2) into a list; 16 bytes per thread
// step, per-byte appends only in waves that found one.
__global__ void compact_unknown_kernel(const uint8_t* __restrict__ sign, int64_t n,
                                       int64_t* __restrict__ list, int64_t cap,
                                       unsigned long long* counter) {
    const int64_t n16 = n >> 4;
    const uint4* s16 = reinterpret_cast<const uint4*>(sign);
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t total_iters = (n16 + stride - 1) / stride;
    int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t it = 0; it < total_iters; ++it) {
        int64_t i = tid + it * stride;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (i < n16) v = s16[i];
        // byte values are 0, 1 or 2: bit 1 marks a 2
        bool any = ((v.x | v.y | v.z | v.w) & 0x02020202u) != 0;
        if (__ballot(any) == 0) continue;
        unsigned words[4] = {v.x, v.y, v.z, v.w};
        for (int b = 0; b < 16; ++b) {
            bool want = ((words[b >> 2] >> (8 * (b & 3))) & 0x2u) != 0;
            long long slot = wave_append(counter, want);
            if (want && slot < cap) list[slot] = i * 16 + b;
        }
    }
    // tail (fewer than 16 bytes), appended by the first lanes of the grid
    int64_t t = (n16 << 4) + tid;
    bool want = (t < n) && (sign[t] == 2);
    long long slot = wave_append(counter, want);
    if (want && slot < cap) list[slot] = t;
}

// [E2] one block per site marks its (2r+1)^3 box (cuda_rasterizer.cu:615-646).
__global__ void mark_patch_kernel(uint8_t* __restrict__ sign, const int64_t* __restrict__ sites,
                                  int n_sites, int r, int dx, int dy, int dz) {
    int s = blockIdx.x;
    if (s >= n_sites) return;
    int64_t v = sites[s];
    int cx = (int)(v % dx), cy = (int)((v / dx) % dy), cz = (int)(v / ((int64_t)dx * dy));
    int x0 = max(cx - r, 0), y0 = max(cy - r, 0), z0 = max(cz - r, 0);
    int x1 = min(cx + r, dx - 1), y1 = min(cy + r, dy - 1), z1 = min(cz + r, dz - 1);
    int64_t nx = x1 - x0 + 1, ny = y1 - y0 + 1;
    int64_t box = nx * ny * (int64_t)(z1 - z0 + 1);
    for (int64_t k = threadIdx.x; k < box; k += blockDim.x) {
        int tz = (int)(k / (nx * ny));
        int64_t rem = k - (int64_t)tz * nx * ny;
        int ty = (int)(rem / nx);
        int tx = (int)(rem - (int64_t)ty * nx);
        sign[(int64_t)(z0 + tz) * dx * dy + (int64_t)(y0 + ty) * dx + (x0 + tx)] = 2;
    }
}

// Stratified probe voxels of generation g, binary64 as the reference
// (cuda_rasterizer.cu:582-586).
__global__ void verify_indices_kernel(int64_t total, int samples, int gen, int64_t* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= samples) return;
    double off = (gen == 0) ? 0.5 : (gen == 1) ? (1.0 / 6.0) : (5.0 / 6.0);
    uint64_t v = (uint64_t)(((double)i + off) * (double)total / (double)samples);
    if (v >= (uint64_t)total) v = (uint64_t)total - 1;
    out[i] = (int64_t)v;
}

__global__ void gather_u8_kernel(const uint8_t* __restrict__ src, const int64_t* __restrict__ idx,
                                 int64_t n, uint8_t* __restrict__ dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void gather_f32_kernel(const float* __restrict__ src, const int64_t* __restrict__ idx, int64_t n,
                                  float* __restrict__ dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void scatter_decision_kernel(const double* __restrict__ w, const int64_t* __restrict__ idx,
                                        int64_t n, uint8_t* __restrict__ sign) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        sign[idx[i]] = (fabs(w[i]) >= 0.5) ? 1 : 0;
}

__global__ void range_decision_kernel(const double* __restrict__ w, int64_t base, int64_t n,
                                      uint8_t* __restrict__ sign) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        sign[base + i] = (fabs(w[i]) >= 0.5) ? 1 : 0;
}

__global__ void compare_samples_kernel(const double* __restrict__ w, const int64_t* __restrict__ idx,
                                       int64_t n, const uint8_t* __restrict__ sign,
                                       int64_t* __restrict__ sites, int64_t cap,
                                       unsigned long long* counter) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool mism = false;
    int64_t v = 0;
    if (i < n) {
        v = idx[i];
        uint8_t s = sign[v];
        bool inside = fabs(w[i]) >= 0.5;
        mism = (s != 2) && ((s == 1) != inside);
    }
    long long slot = wave_append(counter, mism);
    if (mism && slot < cap) sites[slot] = v;
}

// ---- exclusive scan (uint32 counts -> int64 starts) ------------------------
constexpr int kScanBlock = 1024;
constexpr int kScanPerThread = 4;
constexpr int kScanTile = kScanBlock * kScanPerThread;

__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* sh, int64_t* total) {
    // wave inclusive scan
    int lane = threadIdx.x & (kWave - 1);
    int wid = threadIdx.x / kWave;
    int64_t x = v;
    for (int off = 1; off < kWave; off <<= 1) {
        int64_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) sh[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int nw = blockDim.x / kWave;
        int64_t s = (lane < nw) ? sh[lane] : 0;
        for (int off = 1; off < kWave; off <<= 1) {
            int64_t y = __shfl_up(s, off);
            if (lane >= off) s += y;
        }
        if (lane < nw) sh[lane] = s;
    }
    __syncthreads();
    int64_t wave_off = (wid > 0) ? sh[wid - 1] : 0;
    *total = sh[blockDim.x / kWave - 1];
    __syncthreads();
    return wave_off + x - v;
}

template <typename T>
__global__ void scan_tile_kernel(const T* __restrict__ in, int64_t n, int64_t* __restrict__ out,
                                 int64_t* __restrict__ tile_sums) {
    __shared__ int64_t sh[kScanBlock / kWave];
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
    int64_t vals[kScanPerThread];
    int64_t sum = 0;
    for (int k = 0; k < kScanPerThread; ++k) {
        int64_t i = base + k;
        vals[k] = (i < n) ? (int64_t)in[i] : 0;
        sum += vals[k];
    }
    int64_t total;
    int64_t pre = block_exclusive_scan(sum, sh, &total);
    for (int k = 0; k < kScanPerThread; ++k) {
        int64_t i = base + k;
        if (i < n) out[i] = pre;
        pre += vals[k];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void scan_add_kernel(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ offs) {
    int64_t i = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
    int64_t o = offs[blockIdx.x];
    for (int k = 0; k < kScanPerThread; ++k, i += kScanBlock)
        if (i < n) out[i] += o;
}

__global__ void scan_finish_kernel(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ last_in_pre,
                                   const uint32_t* __restrict__ in) {
    // out[n] = out[n-1] + in[n-1]
    if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = (n > 0) ? out[n - 1] + (int64_t)in[n - 1] : 0;
    (void)last_in_pre;
}

hipError_t scan_i64_inplace_sums(int64_t* data, int64_t n, int64_t* scratch, hipStream_t s);

// Exclusive scan of int64 values in place (used for tile sums).
hipError_t scan_i64_inplace_sums(int64_t* data, int64_t n, int64_t* scratch, hipStream_t s) {
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    int64_t* sums = scratch;
    scan_tile_kernel<int64_t><<<(unsigned)tiles, kScanBlock, 0, s>>>(data, n, data, sums);
    if (tiles > 1) {
        hipError_t e = scan_i64_inplace_sums(sums, tiles, scratch + tiles, s);
        if (e != hipSuccess) return e;
        scan_add_kernel<<<(unsigned)tiles, kScanBlock, 0, s>>>(data, n, sums);
    }
    return hipGetLastError();
}

}  // namespace

// ---- launchers ---------------------------------------------------------------
hipError_t launch_zero_i32(int* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    zero_i32_kernel<<<(unsigned)((n + 63) / 64), 64, 0, s>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_fill(float* d, int64_t n, float v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    fill_kernel<<<stream_blocks(n / 4 + 1), kBlock, 0, s>>>(d, n, v);
    return hipGetLastError();
}

hipError_t launch_apply(float* d, const uint8_t* sign, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    apply_kernel<<<stream_blocks(n / 4 + 1), kBlock, 0, s>>>(d, sign, n);
    return hipGetLastError();
}

hipError_t launch_clamp_far(float* d, const uint8_t* sign, int64_t n, float bw, int intended,
                            hipStream_t s) {
    if (n <= 0) return hipSuccess;
    clamp_far_kernel<<<stream_blocks(n), kBlock, 0, s>>>(d, sign, n, bw, intended);
    return hipGetLastError();
}

hipError_t launch_copy_rows(const float* src, int64_t sp, int64_t splane, float* dst, int64_t dp, int64_t dplane,
                            int len, int rpp, int planes, hipStream_t s) {
    if (len <= 0 || rpp <= 0 || planes <= 0) return hipSuccess;
    if (sp < len || dp < len || splane < sp * rpp || dplane < dp * rpp) return hipErrorInvalidValue;
    const int64_t rows = (int64_t)rpp * planes;
    const unsigned blocks = (unsigned)std::min<int64_t>(rows, 16384);
    copy_rows_kernel<<<blocks, kBlock, 0, s>>>(src, sp, splane, dst, dp, dplane, len, rpp, rows);
    return hipGetLastError();
}

hipError_t launch_copy_apply_rows(const float* src, int64_t sp, int64_t splane, float* dst, const uint8_t* sign,
                                  int len, int rpp, int planes, hipStream_t s) {
    if (len <= 0 || rpp <= 0 || planes <= 0) return hipSuccess;
    if (sp < len || splane < sp * rpp) return hipErrorInvalidValue;
    const int64_t rows = (int64_t)rpp * planes;
    const unsigned blocks = (unsigned)std::min<int64_t>(rows, 16384);
    copy_apply_rows_kernel<<<blocks, kBlock, 0, s>>>(src, sp, splane, dst, sign, len, rpp, rows);
    return hipGetLastError();
}

hipError_t launch_count_far_rows(const float* d, int64_t p, int64_t plane, int len, int rpp, int planes,
                                 unsigned long long* counter, hipStream_t s) {
    if (len <= 0 || rpp <= 0 || planes <= 0) return hipSuccess;
    if (p < len || plane < p * rpp) return hipErrorInvalidValue;
    const int64_t rows = (int64_t)rpp * planes;
    const unsigned blocks = (unsigned)std::min<int64_t>(rows, 16384);
    count_far_rows_kernel<<<blocks, kBlock, 0, s>>>(d, p, plane, len, rpp, rows, counter);
    return hipGetLastError();
}

hipError_t launch_count_far(const float* d, int64_t n, unsigned long long* counter, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    count_far_kernel<<<stream_blocks(n / 4 + 1), kBlock, 0, s>>>(d, n, counter);
    return hipGetLastError();
}

hipError_t launch_compact_unknown(const uint8_t* sign, int64_t n, int64_t* list, int64_t cap,
                                  unsigned long long* counter, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    compact_unknown_kernel<<<stream_blocks(n / 16 + 1), kBlock, 0, s>>>(sign, n, list, cap, counter);
    return hipGetLastError();
}

hipError_t launch_mark_patch(uint8_t* sign, const int64_t* sites, int n_sites, int radius,
                             const int32_t dims[3], hipStream_t s) {
    if (n_sites <= 0) return hipSuccess;
    mark_patch_kernel<<<n_sites, kBlock, 0, s>>>(sign, sites, n_sites, radius, dims[0], dims[1], dims[2]);
    return hipGetLastError();
}

hipError_t launch_verify_indices(int64_t total, int samples, int gen, int64_t* out, hipStream_t s) {
    if (samples <= 0) return hipSuccess;
    verify_indices_kernel<<<(samples + kBlock - 1) / kBlock, kBlock, 0, s>>>(total, samples, gen, out);
    return hipGetLastError();
}

hipError_t launch_gather_u8(const uint8_t* src, const int64_t* idx, int64_t n, uint8_t* dst,
                            hipStream_t s) {
    if (n <= 0) return hipSuccess;
    gather_u8_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(src, idx, n, dst);
    return hipGetLastError();
}

hipError_t launch_gather_f32(const float* src, const int64_t* idx, int64_t n, float* dst, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    gather_f32_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(src, idx, n, dst);
    return hipGetLastError();
}

hipError_t launch_scatter_decision(const double* w, const int64_t* idx, int64_t n, uint8_t* sign,
                                   hipStream_t s) {
    if (n <= 0) return hipSuccess;
    scatter_decision_kernel<<<stream_blocks(n), kBlock, 0, s>>>(w, idx, n, sign);
    return hipGetLastError();
}

hipError_t launch_range_decision(const double* w, int64_t base, int64_t n, uint8_t* sign, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    range_decision_kernel<<<stream_blocks(n), kBlock, 0, s>>>(w, base, n, sign);
    return hipGetLastError();
}

hipError_t launch_compare_samples(const double* w, const int64_t* idx, int64_t n,
                                  const uint8_t* sign, int64_t* sites, int64_t cap,
                                  unsigned long long* counter, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    compare_samples_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(w, idx, n, sign, sites,
                                                                                    cap, counter);
    return hipGetLastError();
}

size_t scan_scratch_bytes(int64_t n) {
    // tile sums of every recursion level
    int64_t total = 0, m = n;
    do {
        m = (m + kScanTile - 1) / kScanTile;
        total += m + 1;
    } while (m > 1);
    return (size_t)(total + 8) * sizeof(int64_t);
}

hipError_t launch_scan_u32(const uint32_t* in, int64_t n, int64_t* out, void* scratch, hipStream_t s) {
    if (n <= 0) {
        return hipMemsetAsync(out, 0, sizeof(int64_t), s);
    }
    int64_t tiles = (n + kScanTile - 1) / kScanTile;
    int64_t* sums = reinterpret_cast<int64_t*>(scratch);
    scan_tile_kernel<uint32_t><<<(unsigned)tiles, kScanBlock, 0, s>>>(in, n, out, sums);
    if (tiles > 1) {
        hipError_t e = scan_i64_inplace_sums(sums, tiles, sums + tiles, s);
        if (e != hipSuccess) return e;
        scan_add_kernel<<<(unsigned)tiles, kScanBlock, 0, s>>>(out, n, sums);
    }
    scan_finish_kernel<<<1, 64, 0, s>>>(out, n, nullptr, in);
    return hipGetLastError();
}

}  // namespace rhz
