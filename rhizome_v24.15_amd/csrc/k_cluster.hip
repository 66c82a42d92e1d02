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
// two terms and the bound; the others are summed triangle by triangle with
// the exact kernel's solid-angle routine.  Two levels: the cells' 2 x 2 x 2
// parents are clusters too (the expansion moved to the parent's centre, the
// radius enclosing every child ball), so a probe evaluates every parent and
// descends only into the parents near it.  When the bound cannot separate |w| from 1/2 (plus a
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

// Parent-major id of the cell holding (x, y, z) (clamped into the grid).
__device__ __forceinline__ int64_t cell_id(float x, float y, float z, const ClusterGrid& g) {
    int ix = (int)floorf((x - g.ox) * g.inv_s), iy = (int)floorf((y - g.oy) * g.inv_s),
        iz = (int)floorf((z - g.oz) * g.inv_s);
    ix = min(max(ix, 0), g.nx - 1);
    iy = min(max(iy, 0), g.ny - 1);
    iz = min(max(iz, 0), g.nz - 1);
    const int64_t parent = ((int64_t)(iz >> 1) * g.py() + (iy >> 1)) * g.px() + (ix >> 1);
    return 8 * parent + (ix & 1) + 2 * (iy & 1) + 4 * (iz & 1);
}

__device__ __forceinline__ int64_t cell_of(const Tri9& v, const ClusterGrid& g) {
    const float cx = (v.ax + v.bx + v.cx) * (1.0f / 3.0f);
    const float cy = (v.ay + v.by + v.cy) * (1.0f / 3.0f);
    const float cz = (v.az + v.bz + v.cz) * (1.0f / 3.0f);
    return cell_id(cx, cy, cz, g);
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

// flags[c] = 1 for a non-empty cell (their prefix sums are the cluster slots).
__global__ void nonempty_kernel(const int64_t* __restrict__ start, int64_t ncells, uint32_t* __restrict__ flags) {
    for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < ncells; c += (int64_t)gridDim.x * kBlock)
        flags[c] = start[c + 1] > start[c] ? 1u : 0u;
}

// One thread per cell: centre (vertex-box centre), radius, area vector and
// area of every non-empty cell, written at its slot (cells in id order).
__global__ void cluster_props_kernel(const float* __restrict__ tris, int64_t ncells,
                                     const int64_t* __restrict__ start, const int32_t* __restrict__ ids,
                                     const int64_t* __restrict__ slots, float4* __restrict__ info,
                                     int2* __restrict__ range) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t iters = (ncells + stride - 1) / stride;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + it * stride;
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
        if (want) {
            const int64_t slot = slots[c];
            info[kInfo * slot] = i0;
            info[kInfo * slot + 1] = i1;
            info[kInfo * slot + 2] = i2;
            info[kInfo * slot + 3] = i3;
            info[kInfo * slot + 4] = i4;
            range[slot] = make_int2((int)b, (int)(e - b));
        }
    }
}


// Parent clusters: one thread per parent cell; its children are the clusters
// at slots [slot[8 P], slot[8 P + 8]).  Centre: middle of the box of the
// children's balls; radius R = max_k |c_k - C| + r_k (every triangle of the
// parent lies within R of C); N and A summed; the moment moved to the new
// centre, M' = sum_k M_k + (c_k - C) N_k^T (tr M' = sum_k tr M_k + (c_k - C).N_k).
__global__ void parent_flags_kernel(const int64_t* __restrict__ slots, int64_t npar, uint32_t* __restrict__ flags) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npar; q += (int64_t)gridDim.x * kBlock)
        flags[q] = slots[8 * q + 8] > slots[8 * q] ? 1u : 0u;
}

__global__ void parent_props_kernel(const int64_t* __restrict__ slots, int64_t npar, const float4* __restrict__ info,
                                    const int64_t* __restrict__ pslots, float4* __restrict__ pinfo,
                                    int2* __restrict__ prange) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npar; q += (int64_t)gridDim.x * kBlock) {
        const int64_t b = slots[8 * q], e = slots[8 * q + 8];
        if (e <= b) continue;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t k = b; k < e; ++k) {
            const float4 a = info[kInfo * k];
            const double c[3] = {a.x, a.y, a.z};
            for (int u = 0; u < 3; ++u) {
                lo[u] = fmin(lo[u], c[u] - a.w);
                hi[u] = fmax(hi[u], c[u] + a.w);
            }
        }
        const double C[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
        const float Cf[3] = {(float)C[0], (float)C[1], (float)C[2]};
        double R = 0.0, N[3] = {0, 0, 0}, A = 0.0, M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t k = b; k < e; ++k) {
            const float4 a = info[kInfo * k], n = info[kInfo * k + 1], m0 = info[kInfo * k + 2],
                         m1 = info[kInfo * k + 3], m2 = info[kInfo * k + 4];
            // offsets from the centre as stored (binary32), so the shift is exact for that centre
            const double d[3] = {(double)a.x - Cf[0], (double)a.y - Cf[1], (double)a.z - Cf[2]};
            R = fmax(R, sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + a.w);
            const double nk[3] = {n.x, n.y, n.z};
            const double mk[9] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w, m2.x};
            for (int u = 0; u < 3; ++u) {
                N[u] += nk[u];
                for (int v = 0; v < 3; ++v) M[3 * u + v] += mk[3 * u + v] + d[u] * nk[v];
            }
            A += n.w;
        }
        const int64_t slot = pslots[q];
        pinfo[kInfo * slot] = make_float4(Cf[0], Cf[1], Cf[2], (float)(R * (1.0 + 1e-6)) + 1e-6f);
        pinfo[kInfo * slot + 1] = make_float4((float)N[0], (float)N[1], (float)N[2], (float)(A * (1.0 + 1e-6)));
        pinfo[kInfo * slot + 2] = make_float4((float)M[0], (float)M[1], (float)M[2], (float)M[3]);
        pinfo[kInfo * slot + 3] = make_float4((float)M[4], (float)M[5], (float)M[6], (float)M[7]);
        pinfo[kInfo * slot + 4] = make_float4((float)M[8], (float)(M[0] + M[4] + M[8]), 0.f, 0.f);
        prange[slot] = make_int2((int)b, (int)(e - b));
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
// term + bound (binary32 with hardware reciprocal square root, compensated
// (Kahan) binary32 sum; the bound is inflated by 1% and a 2e-3 margin covers
// the rounding of both); near clusters: exact solid angles of their triangles
// (wave-uniform loop; lanes for which the cluster is far sit it out).  Only
// the finalize step, which adds a probe's partial rows and checks the sum
// against its bound, works in binary64.
// Orders a wave's LDS writes before its later reads (and the reverse).
__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Terms of one cluster (info at ci) for a lane whose position is (px, py,
// pz): far (d > beta r) adds the two-term expansion to the compensated sum
// (far, farc) and its bound to `err` and returns false; near returns true and
// adds nothing.
__device__ __forceinline__ bool cluster_term(const float4* __restrict__ ci, float px, float py, float pz,
                                             float beta2, bool live, float& far, float& farc, float& err) {
    const float4 a = ci[0];
    const float dx = a.x - px, dy = a.y - py, dz = a.z - pz;
    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    const float inv = __builtin_amdgcn_rsqf(d2);
    const float d = d2 * inv;
    const bool near = !(d2 > beta2 * a.w * a.w) || !(d > a.w + 1e-5f);
    if (live && !near) {
        const float4 b = ci[1], m0 = ci[2], m1 = ci[3], m2 = ci[4];
        const float inv3 = inv * inv * inv;
        const float dip = __builtin_fmaf(b.z, dz, __builtin_fmaf(b.y, dy, b.x * dx));
        const float hx = dx * inv, hy = dy * inv, hz = dz * inv;
        // x^T M x / |x|^2 with M row-major (m0.x m0.y m0.z / m0.w m1.x m1.y / m1.z m1.w m2.x)
        const float mx = __builtin_fmaf(m0.z, hz, __builtin_fmaf(m0.y, hy, m0.x * hx));
        const float my = __builtin_fmaf(m1.y, hz, __builtin_fmaf(m1.x, hy, m0.w * hx));
        const float mz = __builtin_fmaf(m2.x, hz, __builtin_fmaf(m1.w, hy, m1.z * hx));
        const float xmx = __builtin_fmaf(hz, mz, __builtin_fmaf(hy, my, hx * mx));
        const float y = (dip + __builtin_fmaf(-3.0f, xmx, m2.y)) * inv3 - farc;
        const float t = far + y;
        farc = (t - far) - y;
        far = t;
        // remainder bound 3 r^2 A / (d - r)^4; the hardware reciprocal
        // (within 1 ulp) is covered by the bound's 1% inflation
        const float g = d - a.w;
        const float g2 = g * g;
        err += 3.0f * a.w * a.w * b.w * __builtin_amdgcn_rcpf(g2 * g2);
    }
    return near;
}

__global__ __launch_bounds__(kBlock) void winding_far_kernel(const float* __restrict__ tris,
                                                             const float4* __restrict__ info,
                                                             const int2* __restrict__ range,
                                                             const int32_t* __restrict__ ids,
                                                             const float4* __restrict__ pinfo,
                                                             const int2* __restrict__ prange, int npar,
                                                             int use_parents, int per_part, Pts P, int64_t n,
                                                             float beta,
                                                             double* __restrict__ part_w,
                                                             float* __restrict__ part_err) {
    // blockIdx.y selects the parent range [per_part y, per_part (y + 1)); the
    // partial sums go to row y of the part arrays, at the sorted position.
    // A parent far from a lane enters through its own expansion; for the
    // lanes it is near, its children do (expansion, or exact triangles when
    // the child is near as well).  With use_parents = 0 every probe descends
    // into every parent: the single-level certificate, with tighter bounds.
    __shared__ float stage[kBlock / 64][9][64];  // per wave: up to 64 triangles, component-major
    const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
    const int64_t k0 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = k0 < n;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (live) voxel_pos(P, P.idx[P.order[k0]], px, py, pz);
    float sum = 0.0f, comp = 0.0f;  // near field, compensated
    float far = 0.0f, farc = 0.0f;  // far field, compensated
    float err = 0.0f;
    const float beta2 = beta * beta;
    const int qb = (int)blockIdx.y * per_part;
    const int qe = min(npar, qb + per_part);
    for (int qc = qb; qc < qe; qc += kCl) {
        bool nearp[kCl];
#pragma unroll
        for (int u = 0; u < kCl; ++u) {
            const int q = qc + u;
            nearp[u] = false;
            if (q >= qe) continue;
            // without parents (the second, tighter pass) every parent is "near"
            nearp[u] = !use_parents || cluster_term(pinfo + kInfo * q, px, py, pz, beta2, live, far, farc, err);
        }
#pragma unroll
        for (int u = 0; u < kCl; ++u) {
            const int q = qc + u;
            if (q >= qe) continue;
            const bool want = live && nearp[u];
            if (__ballot(want) == 0) continue;
            const int2 ch = prange[q];
            for (int cb = ch.x; cb < ch.x + ch.y; cb += kCl) {
                // children in groups (their data loads issue together)
                bool nearc[kCl];
#pragma unroll
                for (int v = 0; v < kCl; ++v) {
                    const int c = cb + v;
                    nearc[v] = false;
                    if (c >= ch.x + ch.y) continue;
                    nearc[v] = cluster_term(info + kInfo * c, px, py, pz, beta2, want, far, farc, err);
                }
#pragma unroll
                for (int v = 0; v < kCl; ++v) {
                    const int c = cb + v;
                    if (c >= ch.x + ch.y) continue;
                    const bool tri = want && nearc[v];
                    if (__ballot(tri) == 0) continue;
                    const int2 rg = range[c];
                    for (int j0 = 0; j0 < rg.y; j0 += 64) {
                        // the wave stages up to 64 of the child's triangles in LDS
                        const int cnt = min(64, rg.y - j0);
                        wave_fence();
                        if (lane < cnt) {
                            const Tri9 t = load_tri(tris, ids[rg.x + j0 + lane]);
                            float* o = &stage[wv][0][lane];
                            o[0 * 64] = t.ax; o[1 * 64] = t.ay; o[2 * 64] = t.az;
                            o[3 * 64] = t.bx; o[4 * 64] = t.by; o[5 * 64] = t.bz;
                            o[6 * 64] = t.cx; o[7 * 64] = t.cy; o[8 * 64] = t.cz;
                        }
                        wave_fence();
                        if (tri) {
                            for (int j = 0; j < cnt; ++j) {
                                Tri9 t;
                                t.ax = stage[wv][0][j]; t.ay = stage[wv][1][j]; t.az = stage[wv][2][j];
                                t.bx = stage[wv][3][j]; t.by = stage[wv][4][j]; t.bz = stage[wv][5][j];
                                t.cx = stage[wv][6][j]; t.cy = stage[wv][7][j]; t.cz = stage[wv][8][j];
                                const float o = solid_angle_fast(px, py, pz, t);
                                const float y = o - comp;
                                const float vv = sum + y;
                                comp = (vv - sum) - y;
                                sum = vv;
                            }
                        }
                    }
                }
            }
        }
    }
    if (live) {
        const int64_t at = (int64_t)blockIdx.y * n + k0;
        part_w[at] = ((double)sum - (double)comp) + ((double)far - (double)farc);
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
    return cell_id(px, py, pz, g);
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
                                int2* range, int64_t* slot, uint32_t* pflags, int64_t* pslot, float4* pinfo,
                                int2* prange, hipStream_t s) {
    const int64_t ncells = g.cells(), npar = g.parents();
    const unsigned tb = (unsigned)std::min<int64_t>((nt + kBlock - 1) / kBlock, 4096);
    cluster_count_kernel<<<tb, kBlock, 0, s>>>(tris, nt, g, counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = launch_scan_u32(counts, ncells, start, scan_scratch, s);
    if (e != hipSuccess) return e;
    cluster_fill_kernel<<<tb, kBlock, 0, s>>>(tris, nt, g, start, cursor, ids);
    const unsigned cb = (unsigned)std::min<int64_t>((ncells + kBlock - 1) / kBlock, 8192);
    nonempty_kernel<<<cb, kBlock, 0, s>>>(start, ncells, counts);  // counts is free again
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_scan_u32(counts, ncells, slot, scan_scratch, s);
    if (e != hipSuccess) return e;
    // one thread per cell with a serial loop over its triangles: small blocks
    // spread the cells over every CU
    const unsigned pcb = (unsigned)std::min<int64_t>((ncells + 63) / 64, 65536);
    cluster_props_kernel<<<pcb, 64, 0, s>>>(tris, ncells, start, ids, slot, info, range);
    const unsigned pb = (unsigned)std::min<int64_t>((npar + kBlock - 1) / kBlock, 8192);
    parent_flags_kernel<<<pb, kBlock, 0, s>>>(slot, npar, pflags);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_scan_u32(pflags, npar, pslot, scan_scratch, s);
    if (e != hipSuccess) return e;
    parent_props_kernel<<<pb, kBlock, 0, s>>>(slot, npar, info, pslot, pinfo, prange);
    return hipGetLastError();
}

hipError_t launch_point_sort(const Grid& g, const int64_t* idx, const int64_t* pos, int64_t m, const ClusterGrid& cg,
                             uint32_t* counts, uint32_t* cursor, int64_t* start, void* scan_scratch, int64_t* order,
                             hipStream_t s) {
    if (m <= 0) return hipSuccess;
    Pts P{idx, nullptr, g.mn[0], g.mn[1], g.mn[2], g.dims[0], g.dims[1], g.h};
    const int64_t ncells = cg.cells();
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
    // parents per part
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
                              const float4* pinfo, const int2* prange, int npar, bool use_parents, const Grid& g,
                              const int64_t* idx, const int64_t* order, int64_t n, float beta, void* scratch,
                              double* w, int64_t* undecided, unsigned long long* n_undecided, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    Pts P{idx, order, g.mn[0], g.mn[1], g.mn[2], g.dims[0], g.dims[1], g.h};
    const int parts = far_parts(n, npar);
    int per_part = (npar + parts - 1) / parts;
    per_part = (per_part + kCl - 1) / kCl * kCl;
    double* part_w = (double*)scratch;
    float* part_err = (float*)(part_w + (size_t)parts * (size_t)n);
    const unsigned bx = (unsigned)((n + kBlock - 1) / kBlock);
    winding_far_kernel<<<dim3(bx, (unsigned)parts), kBlock, 0, s>>>(tris, info, range, ids, pinfo, prange, npar,
                                                                   use_parents ? 1 : 0, per_part, P, n, beta,
                                                                   part_w, part_err);
    winding_far_finalize_kernel<<<bx, kBlock, 0, s>>>(part_w, part_err, parts, order, n, w, undecided, n_undecided);
    return hipGetLastError();
}

}  // namespace rhz
