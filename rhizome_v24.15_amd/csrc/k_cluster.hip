// k_cluster.hip -- certified far-field evaluation of the exact winding sign.
//
// The sign of a voxel is |w| >= 1/2 with w = sum_t Omega_t(p) / (4 pi).  The
// triangles are grouped into spatial clusters (cells of a coarse grid, by
// centroid).  For a cluster C with centre c and radius r (every vertex
// within r of c) and a point p at distance d = |c - p| > r,
//   Omega_t(p) = int_t n_t . f(q - p) dA,   f(x) = x / |x|^3,
// f = -grad(1/|x|), and the k-th directional derivative of 1/|x| is at most
// k! / |x|^(k+1) (Legendre generating function), so |D^2 f(y)| <= 6 / |y|^4.
// The first-order expansion of f around c, integrated over each triangle
// (whose centroid g_t integrates q - c exactly), gives
//   sum_t Omega_t(p) = N.x/|x|^3 + (tr M - 3 x^T M x / |x|^2) / |x|^3 + R,
//   |R| <= 3 r^2 A_C / (d - r)^4,            x = c - p,
// with N = sum_t s_t (s_t the area vector of t), M = sum_t (g_t - c) s_t^T and
// A_C = sum_t |s_t|.  Clusters far from p (d > beta r) enter through these
// two terms and the bound;
// the others are summed triangle by triangle with the exact kernel's
// solid-angle routine.  When the bound cannot separate |w| from 1/2 (plus a
// margin for rounding), the voxel goes to the brute-force exact kernel.  So
// every decision equals the exact kernel's: the certificate only spares the
// work of summing far triangles one by one.
#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

constexpr int kBlock = 256;
constexpr int kInfo = 5;  // float4 per cluster: centre+r, N+A, M (9), tr M
constexpr int kCl = 4;    // clusters whose data a lane loads together

__device__ __forceinline__ int ordered(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}

// Vertex bounding box (ordered-int atomics, order independent) and per-block
// partial sums of triangle area (summed on the host in a fixed order).
__global__ void tri_box_area_kernel(const float* __restrict__ tris, int64_t nt, int* __restrict__ box,
                                    double* __restrict__ area_part) {
    __shared__ double sh[kBlock];
    int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-0x7fffffff - 1, -0x7fffffff - 1, -0x7fffffff - 1};
    double a = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < nt; t += (int64_t)gridDim.x * kBlock) {
        const Tri9 v = load_tri(tris, t);
        const float xs[3] = {v.ax, v.bx, v.cx}, ys[3] = {v.ay, v.by, v.cy}, zs[3] = {v.az, v.bz, v.cz};
        for (int k = 0; k < 3; ++k) {
            lo[0] = min(lo[0], ordered(xs[k]));
            hi[0] = max(hi[0], ordered(xs[k]));
            lo[1] = min(lo[1], ordered(ys[k]));
            hi[1] = max(hi[1], ordered(ys[k]));
            lo[2] = min(lo[2], ordered(zs[k]));
            hi[2] = max(hi[2], ordered(zs[k]));
        }
        const double ux = (double)v.bx - v.ax, uy = (double)v.by - v.ay, uz = (double)v.bz - v.az;
        const double wx = (double)v.cx - v.ax, wy = (double)v.cy - v.ay, wz = (double)v.cz - v.az;
        const double cx = uy * wz - uz * wy, cy = uz * wx - ux * wz, cz = ux * wy - uy * wx;
        a += 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
    }
    for (int k = 0; k < 3; ++k) {
        atomicMin(&box[k], lo[k]);
        atomicMax(&box[3 + k], hi[k]);
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) area_part[blockIdx.x] = sh[0];
}

__device__ __forceinline__ int64_t cell_of(const Tri9& v, const ClusterGrid& g) {
    const float cx = (v.ax + v.bx + v.cx) * (1.0f / 3.0f);
    const float cy = (v.ay + v.by + v.cy) * (1.0f / 3.0f);
    const float cz = (v.az + v.bz + v.cz) * (1.0f / 3.0f);
    int ix = (int)floorf((cx - g.ox) * g.inv_s), iy = (int)floorf((cy - g.oy) * g.inv_s),
        iz = (int)floorf((cz - g.oz) * g.inv_s);
    ix = min(max(ix, 0), g.nx - 1);
    iy = min(max(iy, 0), g.ny - 1);
    iz = min(max(iz, 0), g.nz - 1);
    return ((int64_t)iz * g.ny + iy) * g.nx + ix;
}

__global__ void cluster_count_kernel(const float* __restrict__ tris, int64_t nt, ClusterGrid g,
                                     uint32_t* __restrict__ counts) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < nt; t += (int64_t)gridDim.x * kBlock)
        atomicAdd(&counts[cell_of(load_tri(tris, t), g)], 1u);
}

__global__ void cluster_fill_kernel(const float* __restrict__ tris, int64_t nt, ClusterGrid g,
                                    const int64_t* __restrict__ start, uint32_t* __restrict__ cursor,
                                    int32_t* __restrict__ ids) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < nt; t += (int64_t)gridDim.x * kBlock) {
        const int64_t c = cell_of(load_tri(tris, t), g);
        const uint32_t k = atomicAdd(&cursor[c], 1u);
        ids[start[c] + k] = (int32_t)t;
    }
}

// One thread per cell: centre (vertex-box centre), radius, area vector and
// area of every non-empty cell, appended to the cluster list.
__global__ void cluster_props_kernel(const float* __restrict__ tris, int64_t ncells,
                                     const int64_t* __restrict__ start, const int32_t* __restrict__ ids,
                                     float4* __restrict__ info, int2* __restrict__ range,
                                     unsigned long long* __restrict__ count) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const int64_t iters = (ncells + stride - 1) / stride;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x + it * stride;
        int64_t b = 0, e = 0;
        if (c < ncells) {
            b = start[c];
            e = start[c + 1];
        }
        const bool want = e > b;
        float4 i0 = make_float4(0.f, 0.f, 0.f, 0.f), i1 = i0, i2 = i0, i3 = i0, i4 = i0;
        if (want) {
            float lx = 3e38f, ly = 3e38f, lz = 3e38f, hx = -3e38f, hy = -3e38f, hz = -3e38f;
            for (int64_t k = b; k < e; ++k) {
                const Tri9 v = load_tri(tris, ids[k]);
                lx = fminf(lx, fminf(v.ax, fminf(v.bx, v.cx)));
                ly = fminf(ly, fminf(v.ay, fminf(v.by, v.cy)));
                lz = fminf(lz, fminf(v.az, fminf(v.bz, v.cz)));
                hx = fmaxf(hx, fmaxf(v.ax, fmaxf(v.bx, v.cx)));
                hy = fmaxf(hy, fmaxf(v.ay, fmaxf(v.by, v.cy)));
                hz = fmaxf(hz, fmaxf(v.az, fmaxf(v.bz, v.cz)));
            }
            const float cx = 0.5f * (lx + hx), cy = 0.5f * (ly + hy), cz = 0.5f * (lz + hz);
            double r2 = 0.0, nx = 0.0, ny = 0.0, nz = 0.0, area = 0.0;
            double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int64_t k = b; k < e; ++k) {
                const Tri9 v = load_tri(tris, ids[k]);
                const double px[3] = {v.ax, v.bx, v.cx}, py[3] = {v.ay, v.by, v.cy}, pz[3] = {v.az, v.bz, v.cz};
                for (int j = 0; j < 3; ++j) {
                    const double dx = px[j] - cx, dy = py[j] - cy, dz = pz[j] - cz;
                    r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
                }
                const double ux = px[1] - px[0], uy = py[1] - py[0], uz = pz[1] - pz[0];
                const double wx = px[2] - px[0], wy = py[2] - py[0], wz = pz[2] - pz[0];
                const double ax = 0.5 * (uy * wz - uz * wy), ay = 0.5 * (uz * wx - ux * wz),
                             az = 0.5 * (ux * wy - uy * wx);
                nx += ax;
                ny += ay;
                nz += az;
                area += sqrt(ax * ax + ay * ay + az * az);
                const double g[3] = {(px[0] + px[1] + px[2]) / 3.0 - cx, (py[0] + py[1] + py[2]) / 3.0 - cy,
                                     (pz[0] + pz[1] + pz[2]) / 3.0 - cz};
                const double sv[3] = {ax, ay, az};
                for (int u = 0; u < 3; ++u)
                    for (int q = 0; q < 3; ++q) M[3 * u + q] += g[u] * sv[q];
            }
            // radius and area rounded up: the bound stays an upper bound
            i0 = make_float4(cx, cy, cz, (float)(sqrt(r2) * (1.0 + 1e-6)) + 1e-6f);
            i1 = make_float4((float)nx, (float)ny, (float)nz, (float)(area * (1.0 + 1e-6)));
            i2 = make_float4((float)M[0], (float)M[1], (float)M[2], (float)M[3]);
            i3 = make_float4((float)M[4], (float)M[5], (float)M[6], (float)M[7]);
            i4 = make_float4((float)M[8], (float)(M[0] + M[4] + M[8]), 0.f, 0.f);
        }
        const long long slot = wave_append(count, want);
        if (want) {
            info[kInfo * slot] = i0;
            info[kInfo * slot + 1] = i1;
            info[kInfo * slot + 2] = i2;
            info[kInfo * slot + 3] = i3;
            info[kInfo * slot + 4] = i4;
            range[slot] = make_int2((int)b, (int)(e - b));
        }
    }
}

struct Pts {
    const int64_t* idx;   // voxel ids, indexed by list position
    const int64_t* order; // list positions in the order the lanes take them (spatially sorted)
    int mnx, mny, mnz, dx, dy;
    float h;
};

__device__ __forceinline__ void voxel_pos(const Pts& P, int64_t v, float& px, float& py, float& pz) {
    const int x = (int)(v % P.dx);
    const int64_t rr = v / P.dx;
    px = lattice(P.mnx, x, P.h);
    py = lattice(P.mny, (int)(rr % P.dy), P.h);
    pz = lattice(P.mnz, (int)(rr / P.dy), P.h);
}

// Lane k takes list position order[k]; lanes of a wave are spatially close,
// so the clusters near one lane are near the others.  Far clusters: dipole
// term + bound (binary32 with hardware reciprocal square root, accumulated
// in binary64; the bound is inflated by 1% and a 2e-3 margin covers the
// rounding of both); near clusters: exact solid angles of their triangles
// (wave-uniform loop; lanes for which the cluster is far sit it out).
__global__ __launch_bounds__(kBlock) void winding_far_kernel(const float* __restrict__ tris,
                                                             const float4* __restrict__ info,
                                                             const int2* __restrict__ range,
                                                             const int32_t* __restrict__ ids, int ncl, int per_part,
                                                             Pts P, int64_t n, float beta,
                                                             double* __restrict__ part_w,
                                                             float* __restrict__ part_err) {
    // blockIdx.y selects the cluster range [per_part y, per_part (y + 1));
    // the partial sums go to row y of the part arrays, at the sorted position.
    const int64_t k0 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = k0 < n;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (live) voxel_pos(P, P.idx[P.order[k0]], px, py, pz);
    float sum = 0.0f, comp = 0.0f;  // near field, compensated
    double far = 0.0;
    float err = 0.0f;
    const float beta2 = beta * beta;
    const int kb = (int)blockIdx.y * per_part;
    const int ke = min(ncl, kb + per_part);
    for (int kc = kb; kc < ke; kc += kCl) {
        bool nearv[kCl];
#pragma unroll
        for (int q = 0; q < kCl; ++q) {
            const int k = kc + q;
            nearv[q] = false;
            if (k >= ke) continue;
            const float4 a = info[kInfo * k], b = info[kInfo * k + 1], m0 = info[kInfo * k + 2],
                         m1 = info[kInfo * k + 3], m2 = info[kInfo * k + 4];
            const float dx = a.x - px, dy = a.y - py, dz = a.z - pz;
            const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            const float inv = __builtin_amdgcn_rsqf(d2);
            const float d = d2 * inv;
            const bool near = !(d2 > beta2 * a.w * a.w) || !(d > a.w + 1e-5f);
            nearv[q] = near;
            if (live && !near) {
                const float inv3 = inv * inv * inv;
                const float dip = __builtin_fmaf(b.z, dz, __builtin_fmaf(b.y, dy, b.x * dx));
                const float hx = dx * inv, hy = dy * inv, hz = dz * inv;
                // x^T M x / |x|^2 with M row-major (m0.x m0.y m0.z / m0.w m1.x m1.y / m1.z m1.w m2.x)
                const float mx = __builtin_fmaf(m0.z, hz, __builtin_fmaf(m0.y, hy, m0.x * hx));
                const float my = __builtin_fmaf(m1.y, hz, __builtin_fmaf(m1.x, hy, m0.w * hx));
                const float mz = __builtin_fmaf(m2.x, hz, __builtin_fmaf(m1.w, hy, m1.z * hx));
                const float xmx = __builtin_fmaf(hz, mz, __builtin_fmaf(hy, my, hx * mx));
                far += (double)((dip + __builtin_fmaf(-3.0f, xmx, m2.y)) * inv3);
                // remainder bound 3 r^2 A / (d - r)^4; the hardware reciprocal
                // (within 1 ulp) is covered by the bound's 1% inflation
                const float g = d - a.w;
                const float g2 = g * g;
                err += 3.0f * a.w * a.w * b.w * __builtin_amdgcn_rcpf(g2 * g2);
            }
        }
#pragma unroll
        for (int q = 0; q < kCl; ++q) {
            const int k = kc + q;
            if (k >= ke) continue;
            if (__ballot(live && nearv[q]) != 0) {
                const int2 rg = range[k];
                for (int j = 0; j < rg.y; ++j) {
                    const Tri9 t = load_tri(tris, ids[rg.x + j]);
                    if (live && nearv[q]) {
                        const float o = solid_angle_fast(px, py, pz, t);
                        const float y = o - comp;
                        const float u = sum + y;
                        comp = (u - sum) - y;
                        sum = u;
                    }
                }
            }
        }
    }
    if (live) {
        const int64_t at = (int64_t)blockIdx.y * n + k0;
        part_w[at] = ((double)sum - (double)comp + far);
        part_err[at] = err;
    }
}

// Sum of the partial rows in a fixed order (deterministic), the winding
// number, and the decision: undecided positions are appended to `undecided`.
__global__ __launch_bounds__(kBlock) void winding_far_finalize_kernel(const double* __restrict__ part_w,
                                                                      const float* __restrict__ part_err, int parts,
                                                                      const int64_t* __restrict__ order, int64_t n,
                                                                      double* __restrict__ w,
                                                                      int64_t* __restrict__ undecided,
                                                                      unsigned long long* __restrict__ n_undecided) {
    const int64_t k0 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = k0 < n;
    bool decided = true;
    int64_t i = 0;
    if (live) {
        double sw = 0.0, se = 0.0;
        for (int q = 0; q < parts; ++q) {
            sw += part_w[(int64_t)q * n + k0];
            se += (double)part_err[(int64_t)q * n + k0];
        }
        constexpr double kInv4PiD = 0.07957747154594767;
        const double wv = sw * kInv4PiD;
        const double bound = se * kInv4PiD * 1.01 + 2e-3;
        decided = fabs(fabs(wv) - 0.5) > bound;
        i = order[k0];
        w[i] = wv;
    }
    const long long slot = wave_append(n_undecided, live && !decided);
    if (live && !decided) undecided[slot] = i;
}

// Counting sort of list positions by the cluster-grid cell of their voxel.
__device__ __forceinline__ int64_t point_cell(const Pts& P, int64_t v, const ClusterGrid& g) {
    float px, py, pz;
    voxel_pos(P, v, px, py, pz);
    int ix = (int)floorf((px - g.ox) * g.inv_s), iy = (int)floorf((py - g.oy) * g.inv_s),
        iz = (int)floorf((pz - g.oz) * g.inv_s);
    ix = min(max(ix, 0), g.nx - 1);
    iy = min(max(iy, 0), g.ny - 1);
    iz = min(max(iz, 0), g.nz - 1);
    return ((int64_t)iz * g.ny + iy) * g.nx + ix;
}

__global__ void point_count_kernel(Pts P, const int64_t* __restrict__ pos, int64_t m, ClusterGrid g,
                                   uint32_t* __restrict__ counts) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < m) atomicAdd(&counts[point_cell(P, P.idx[pos ? pos[k] : k], g)], 1u);
}

__global__ void point_fill_kernel(Pts P, const int64_t* __restrict__ pos, int64_t m, ClusterGrid g,
                                  const int64_t* __restrict__ start, uint32_t* __restrict__ cursor,
                                  int64_t* __restrict__ order) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < m) {
        const int64_t i = pos ? pos[k] : k;
        const int64_t c = point_cell(P, P.idx[i], g);
        order[start[c] + atomicAdd(&cursor[c], 1u)] = i;
    }
}

__global__ void gather_i64_kernel(const int64_t* __restrict__ src, const int64_t* __restrict__ pos, int64_t m,
                                  int64_t* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < m) out[k] = src[pos[k]];
}

__global__ void scatter_f64_kernel(const double* __restrict__ src, const int64_t* __restrict__ pos, int64_t m,
                                   double* __restrict__ dst) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < m) dst[pos[k]] = src[k];
}

}  // namespace

hipError_t launch_gather_i64(const int64_t* src, const int64_t* pos, int64_t m, int64_t* out, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    gather_i64_kernel<<<(unsigned)((m + kBlock - 1) / kBlock), kBlock, 0, s>>>(src, pos, m, out);
    return hipGetLastError();
}

hipError_t launch_scatter_f64(const double* src, const int64_t* pos, int64_t m, double* dst, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    scatter_f64_kernel<<<(unsigned)((m + kBlock - 1) / kBlock), kBlock, 0, s>>>(src, pos, m, dst);
    return hipGetLastError();
}

hipError_t launch_tri_box_area(const float* tris, int64_t nt, int* box, double* area_part, int nblocks,
                               hipStream_t s) {
    tri_box_area_kernel<<<(unsigned)nblocks, kBlock, 0, s>>>(tris, nt, box, area_part);
    return hipGetLastError();
}

hipError_t launch_cluster_build(const float* tris, int64_t nt, const ClusterGrid& g, uint32_t* counts,
                                uint32_t* cursor, int64_t* start, void* scan_scratch, int32_t* ids, float4* info,
                                int2* range, unsigned long long* n_clusters, hipStream_t s) {
    const int64_t ncells = (int64_t)g.nx * g.ny * g.nz;
    const unsigned tb = (unsigned)std::min<int64_t>((nt + kBlock - 1) / kBlock, 4096);
    cluster_count_kernel<<<tb, kBlock, 0, s>>>(tris, nt, g, counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_scan_u32(counts, ncells, start, scan_scratch, s);
    if (e != hipSuccess) return e;
    cluster_fill_kernel<<<tb, kBlock, 0, s>>>(tris, nt, g, start, cursor, ids);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const unsigned cb = (unsigned)std::min<int64_t>((ncells + kBlock - 1) / kBlock, 8192);
    cluster_props_kernel<<<cb, kBlock, 0, s>>>(tris, ncells, start, ids, info, range, n_clusters);
    return hipGetLastError();
}

hipError_t launch_point_sort(const Grid& g, const int64_t* idx, const int64_t* pos, int64_t m, const ClusterGrid& cg,
                             uint32_t* counts, uint32_t* cursor, int64_t* start, void* scan_scratch, int64_t* order,
                             hipStream_t s) {
    if (m <= 0) return hipSuccess;
    Pts P{idx, nullptr, g.mn[0], g.mn[1], g.mn[2], g.dims[0], g.dims[1], g.h};
    const int64_t ncells = (int64_t)cg.nx * cg.ny * cg.nz;
    hipError_t e = hipMemsetAsync(counts, 0, sizeof(uint32_t) * (size_t)ncells, s);
    if (e == hipSuccess) e = hipMemsetAsync(cursor, 0, sizeof(uint32_t) * (size_t)ncells, s);
    if (e != hipSuccess) return e;
    const unsigned b = (unsigned)((m + kBlock - 1) / kBlock);
    point_count_kernel<<<b, kBlock, 0, s>>>(P, pos, m, cg, counts);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_scan_u32(counts, ncells, start, scan_scratch, s);
    if (e != hipSuccess) return e;
    point_fill_kernel<<<b, kBlock, 0, s>>>(P, pos, m, cg, start, cursor, order);
    return hipGetLastError();
}

int far_parts(int64_t n, int ncl) {
    // enough blocks to fill the chip (>= 4096) without splitting below kCl
    // clusters per part
    const int64_t bx = (n + kBlock - 1) / kBlock;
    int64_t parts = (4096 + bx - 1) / std::max<int64_t>(bx, 1);
    parts = std::min<int64_t>(parts, std::max(1, (ncl + kCl - 1) / kCl));
    return (int)std::max<int64_t>(1, std::min<int64_t>(parts, 64));
}

size_t far_scratch_bytes(int64_t n, int ncl) {
    const size_t parts = (size_t)far_parts(n, ncl);
    return parts * (size_t)n * (sizeof(double) + sizeof(float)) + 256;
}

hipError_t launch_winding_far(const float* tris, const float4* info, const int2* range, const int32_t* ids,
                              int ncl, const Grid& g, const int64_t* idx, const int64_t* order, int64_t n, float beta,
                              void* scratch, double* w, int64_t* undecided, unsigned long long* n_undecided,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    Pts P{idx, order, g.mn[0], g.mn[1], g.mn[2], g.dims[0], g.dims[1], g.h};
    const int parts = far_parts(n, ncl);
    int per_part = (ncl + parts - 1) / parts;
    per_part = (per_part + kCl - 1) / kCl * kCl;
    double* part_w = (double*)scratch;
    float* part_err = (float*)(part_w + (size_t)parts * (size_t)n);
    const unsigned bx = (unsigned)((n + kBlock - 1) / kBlock);
    winding_far_kernel<<<dim3(bx, (unsigned)parts), kBlock, 0, s>>>(tris, info, range, ids, ncl, per_part, P, n, beta,
                                                                   part_w, part_err);
    winding_far_finalize_kernel<<<bx, kBlock, 0, s>>>(part_w, part_err, parts, order, n, w, undecided, n_undecided);
    return hipGetLastError();
}

}  // namespace rhz
