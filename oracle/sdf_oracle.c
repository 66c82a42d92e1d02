/*
 * sdf_oracle.c -- CPU restatement of the reference Fast SDF path.
 * TEST INFRASTRUCTURE ONLY (see sdf_oracle.h): the checker for the HIP path,
 * never linked into the product.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference repository root.
 */
#include "sdf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <time.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define ORC_INV_4PI_F 0.07957747154594767f
#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* thread pool: parallel_for over [0, n) in chunks                      */
/* ------------------------------------------------------------------ */
static int g_threads = 0;

int orc_num_threads(void) {
    if (g_threads <= 0) {
        const char* e = getenv("RHZ_ORACLE_THREADS");
        long n = e ? strtol(e, NULL, 10) : sysconf(_SC_NPROCESSORS_ONLN);
        g_threads = (n > 0) ? (int)n : 1;
    }
    return g_threads;
}

void orc_set_num_threads(int n) { g_threads = n; }

typedef void (*orc_range_fn)(void* ctx, int64_t begin, int64_t end, int tid);

typedef struct {
    orc_range_fn fn;
    void* ctx;
    int64_t n;
    int64_t chunk;
    atomic_llong next;
    int tid;
} orc_pool_job;

typedef struct {
    orc_pool_job* job;
    int tid;
} orc_worker_arg;

static void* orc_worker(void* p) {
    orc_worker_arg* a = (orc_worker_arg*)p;
    orc_pool_job* j = a->job;
    for (;;) {
        int64_t b = atomic_fetch_add(&j->next, j->chunk);
        if (b >= j->n) break;
        int64_t e = b + j->chunk;
        if (e > j->n) e = j->n;
        j->fn(j->ctx, b, e, a->tid);
    }
    return NULL;
}

static void orc_parallel_for(int64_t n, int64_t chunk, orc_range_fn fn, void* ctx) {
    if (n <= 0) return;
    int nt = orc_num_threads();
    if (chunk < 1) chunk = 1;
    if (nt <= 1 || n <= chunk) {
        fn(ctx, 0, n, 0);
        return;
    }
    orc_pool_job job;
    job.fn = fn;
    job.ctx = ctx;
    job.n = n;
    job.chunk = chunk;
    atomic_init(&job.next, 0);
    pthread_t th[256];
    orc_worker_arg args[256];
    if (nt > 256) nt = 256;
    for (int i = 0; i < nt; ++i) {
        args[i].job = &job;
        args[i].tid = i;
        pthread_create(&th[i], NULL, orc_worker, &args[i]);
    }
    for (int i = 0; i < nt; ++i) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------ */
/* defaults (cuda_rasterizer.cu:74-82, :825, :867, :1010, :1027, :298) */
/* ------------------------------------------------------------------ */
void orc_params_default(orc_params* p) {
    p->fast_min_voxels = 32LL * 1024LL * 1024LL;
    p->band_voxels = 8;
    p->parity_max_hits = 128;
    p->verify_samples = 4096;
    p->verify_gens = 3;
    p->max_sites_per_probe = 64;
    p->patch_radius = 24;
    p->max_rounds = 8;
    p->max_extra_rounds = 16;
    p->sup = 4;
    p->csr_max_entries = 100000000LL;
    p->jitter_x = 0.037f;
    p->jitter_y = 0.051f;
    p->clamp_intended_sign = 1;
    p->fsm_converge = 0;
    p->fsm_round_cap = 512;
    p->winding_mode = 3;
}

/* ------------------------------------------------------------------ */
/* L0 math                                                              */
/* ------------------------------------------------------------------ */
static inline float dotf(const float a[3], const float b[3]) {
    /* glm::dot for vec3: (x*x' + y*y') + z*z' */
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void subf(float o[3], const float a[3], const float b[3]) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}

/* Ericson closest-point regions; math_utils.cuh:43-106. */
float orc_point_triangle_distance_sq(const float p[3], const float a[3],
                                     const float b[3], const float c[3]) {
    float ab[3], ac[3], ap[3];
    subf(ab, b, a);
    subf(ac, c, a);
    subf(ap, p, a);
    float d1 = dotf(ab, ap);
    float d2 = dotf(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return dotf(ap, ap);

    float bp[3];
    subf(bp, p, b);
    float d3 = dotf(ab, bp);
    float d4 = dotf(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return dotf(bp, bp);

    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        float q[3] = {a[0] + v * ab[0], a[1] + v * ab[1], a[2] + v * ab[2]};
        float df[3];
        subf(df, p, q);
        return dotf(df, df);
    }

    float cp[3];
    subf(cp, p, c);
    float d5 = dotf(ab, cp);
    float d6 = dotf(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return dotf(cp, cp);

    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float w = d2 / (d2 - d6);
        float q[3] = {a[0] + w * ac[0], a[1] + w * ac[1], a[2] + w * ac[2]};
        float df[3];
        subf(df, p, q);
        return dotf(df, df);
    }

    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        float cb[3];
        subf(cb, c, b);
        float q[3] = {b[0] + w * cb[0], b[1] + w * cb[1], b[2] + w * cb[2]};
        float df[3];
        subf(df, p, q);
        return dotf(df, df);
    }

    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom;
    float w = vc * denom;
    float q[3] = {a[0] + ab[0] * v + ac[0] * w, a[1] + ab[1] * v + ac[1] * w,
                  a[2] + ab[2] * v + ac[2] * w};
    float df[3];
    subf(df, p, q);
    return dotf(df, df);
}

static inline double dotd(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* Same region logic in binary64 (precision study oracle). */
double orc_point_triangle_distance_sq_f64(const double p[3], const double a[3],
                                          const double b[3], const double c[3]) {
    double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double d1 = dotd(ab, ap), d2 = dotd(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return dotd(ap, ap);
    double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    double d3 = dotd(ab, bp), d4 = dotd(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return dotd(bp, bp);
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        double df[3] = {p[0] - (a[0] + v * ab[0]), p[1] - (a[1] + v * ab[1]),
                        p[2] - (a[2] + v * ab[2])};
        return dotd(df, df);
    }
    double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    double d5 = dotd(ab, cp), d6 = dotd(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return dotd(cp, cp);
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        double df[3] = {p[0] - (a[0] + w * ac[0]), p[1] - (a[1] + w * ac[1]),
                        p[2] - (a[2] + w * ac[2])};
        return dotd(df, df);
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        double df[3] = {p[0] - (b[0] + w * (c[0] - b[0])), p[1] - (b[1] + w * (c[1] - b[1])),
                        p[2] - (b[2] + w * (c[2] - b[2]))};
        return dotd(df, df);
    }
    double den = 1.0 / (va + vb + vc);
    double v = vb * den, w = vc * den;
    double df[3] = {p[0] - (a[0] + ab[0] * v + ac[0] * w), p[1] - (a[1] + ab[1] * v + ac[1] * w),
                    p[2] - (a[2] + ab[2] * v + ac[2] * w)};
    return dotd(df, df);
}

/* Van Oosterom-Strackee solid angle; math_utils.cuh:109-132
 * (vertex guard <= 1e-6 returns 0, :121-123). */
float orc_triangle_solid_angle(const float p[3], const float v0[3],
                               const float v1[3], const float v2[3]) {
    float a[3], b[3], c[3];
    subf(a, v0, p);
    subf(b, v1, p);
    subf(c, v2, p);
    float la = sqrtf(dotf(a, a));
    float lb = sqrtf(dotf(b, b));
    float lc = sqrtf(dotf(c, c));
    if (la <= 1e-6f || lb <= 1e-6f || lc <= 1e-6f) return 0.0f;
    float det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                a[2] * (b[0] * c[1] - b[1] * c[0]);
    float denom = la * lb * lc + dotf(a, b) * lc + dotf(b, c) * la + dotf(c, a) * lb;
    return 2.0f * atan2f(det, denom);
}

double orc_triangle_solid_angle_f64(const double p[3], const double v0[3],
                                    const double v1[3], const double v2[3]) {
    double a[3] = {v0[0] - p[0], v0[1] - p[1], v0[2] - p[2]};
    double b[3] = {v1[0] - p[0], v1[1] - p[1], v1[2] - p[2]};
    double c[3] = {v2[0] - p[0], v2[1] - p[1], v2[2] - p[2]};
    double la = sqrt(dotd(a, a)), lb = sqrt(dotd(b, b)), lc = sqrt(dotd(c, c));
    if (la <= 1e-6 || lb <= 1e-6 || lc <= 1e-6) return 0.0;
    double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                 a[2] * (b[0] * c[1] - b[1] * c[0]);
    double denom = la * lb * lc + dotd(a, b) * lc + dotd(b, c) * la + dotd(c, a) * lb;
    return 2.0 * atan2(det, denom);
}

/* Godunov upwind update; cuda_rasterizer.cu:405-425 (the s1 recovery by
 * sum-minus-extremes is kept as written). */
float orc_godunov(float ax, float ay, float az, float h) {
    float s0 = fminf(ax, fminf(ay, az));
    float s2 = fmaxf(ax, fmaxf(ay, az));
    float s1 = ax + ay + az - s0 - s2;
    float updated = s0 + h;
    if (updated > s1) {
        float disc = 2.0f * h * h - (s1 - s0) * (s1 - s0);
        if (disc >= 0.0f) {
            updated = (s0 + s1 + sqrtf(disc)) * 0.5f;
            if (updated > s2) {
                float disc3 = 3.0f * h * h -
                              ((s1 - s0) * (s1 - s0) + (s2 - s1) * (s2 - s1) +
                               (s2 - s0) * (s2 - s0)) * 0.5f;
                if (disc3 >= 0.0f) updated = (s0 + s1 + s2 + sqrtf(disc3)) / 3.0f;
            }
        }
    }
    return updated;
}

/* ------------------------------------------------------------------ */
/* mesh intake                                                          */
/* ------------------------------------------------------------------ */
/* dev_try_build_cuda_triangle (math_utils.cuh:484-514), compacted in index
 * order (the reference's atomicAdd slot order, cuda_rasterizer.cu:1334, is
 * nondeterministic; results downstream do not depend on it except through
 * the FP32 summation order of the winding number). */
int64_t orc_build_triangles(const float* verts, int64_t nv, const uint32_t* idx,
                            int64_t nt, orc_tri* out) {
    int64_t k = 0;
    for (int64_t t = 0; t < nt; ++t) {
        uint32_t i0 = idx[3 * t], i1 = idx[3 * t + 1], i2 = idx[3 * t + 2];
        if (i0 >= (uint64_t)nv || i1 >= (uint64_t)nv || i2 >= (uint64_t)nv) continue;
        const float* a = verts + 3 * (int64_t)i0;
        const float* b = verts + 3 * (int64_t)i1;
        const float* c = verts + 3 * (int64_t)i2;
        float e1[3], e2[3], n[3];
        subf(e1, b, a);
        subf(e2, c, a);
        /* glm::cross */
        n[0] = e1[1] * e2[2] - e2[1] * e1[2];
        n[1] = e1[2] * e2[0] - e2[2] * e1[0];
        n[2] = e1[0] * e2[1] - e2[0] * e1[1];
        float len2 = dotf(n, n);
        if (len2 > 1e-12f) {
            /* rsqrtf: approximate on the device; the normal is not read by any
               stage of the path, so 1/sqrtf is used here. */
            float r = 1.0f / sqrtf(len2);
            for (int j = 0; j < 3; ++j) {
                out[k].v0[j] = a[j];
                out[k].v1[j] = b[j];
                out[k].v2[j] = c[j];
                out[k].n[j] = n[j] * r;
            }
            ++k;
        }
    }
    return k;
}

/* gpu_reduce_mesh_bbox_kernel (cuda_mesh_ops.cu:71-124): min/max per axis,
 * initial +-1e30. */
void orc_mesh_bbox(const float* verts, int64_t nv, float bmin[3], float bmax[3]) {
    for (int j = 0; j < 3; ++j) {
        bmin[j] = 1e30f;
        bmax[j] = -1e30f;
    }
    for (int64_t i = 0; i < nv; ++i)
        for (int j = 0; j < 3; ++j) {
            float v = verts[3 * i + j];
            if (v < bmin[j]) bmin[j] = v;
            if (v > bmax[j]) bmax[j] = v;
        }
}

/* build_from_mesh framing; cuda_rasterizer.cu:1358-1372. */
void orc_frame_grid(const float* verts, int64_t nv, float h, int32_t min_c[3],
                    int32_t max_c[3]) {
    float bmin[3], bmax[3];
    orc_mesh_bbox(verts, nv, bmin, bmax);
    float pad = h * 10.0f;
    for (int j = 0; j < 3; ++j) {
        float lo = bmin[j] - pad;
        float hi = bmax[j] + pad;
        min_c[j] = (int32_t)floorf(lo / h);
        max_c[j] = (int32_t)ceilf(hi / h);
    }
}

/* ------------------------------------------------------------------ */
/* stages                                                               */
/* ------------------------------------------------------------------ */
/* gpu_fill_float_kernel; cuda_rasterizer.cu:194-203. */
typedef struct {
    float* d;
    const uint8_t* sign;
    float v;
    int intended;
    atomic_llong count;
} elem_ctx;

static void fill_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    elem_ctx* c = (elem_ctx*)vctx;
    for (int64_t i = b; i < e; ++i) c->d[i] = c->v;
}

void orc_fill(float* d, int64_t n, float v) {
    elem_ctx c;
    c.d = d;
    c.v = v;
    orc_parallel_for(n, 1 << 20, fill_range, &c);
}

typedef struct {
    const orc_tri* t;
    int64_t nt;
    float h, bw;
    const int32_t* mn;
    const int32_t* dm;
    float* d;
} splat_ctx;

static inline float tri_min(float a, float b, float c) { return fminf(a, fminf(b, c)); }
static inline float tri_max(float a, float b, float c) { return fmaxf(a, fmaxf(b, c)); }

/* One z-layer of gpu_splat_band_distances_kernel (cuda_rasterizer.cu:211-268)
 * for every triangle; z-layers are disjoint so threads never share a voxel. */
static void splat_range(void* vctx, int64_t zb, int64_t ze, int tid) {
    (void)tid;
    splat_ctx* c = (splat_ctx*)vctx;
    const float h = c->h, bw = c->bw, band_sq = bw * bw;
    const int dx = c->dm[0], dy = c->dm[1];
    const size_t zstride = (size_t)dy * (size_t)dx;
    for (int64_t ti = 0; ti < c->nt; ++ti) {
        const orc_tri* tr = &c->t[ti];
        float tx0 = tri_min(tr->v0[0], tr->v1[0], tr->v2[0]);
        float tx1 = tri_max(tr->v0[0], tr->v1[0], tr->v2[0]);
        float ty0 = tri_min(tr->v0[1], tr->v1[1], tr->v2[1]);
        float ty1 = tri_max(tr->v0[1], tr->v1[1], tr->v2[1]);
        float tz0 = tri_min(tr->v0[2], tr->v1[2], tr->v2[2]);
        float tz1 = tri_max(tr->v0[2], tr->v1[2], tr->v2[2]);
        int x0 = (int)ceilf((tx0 - bw) / h) - c->mn[0];
        int x1 = (int)floorf((tx1 + bw) / h) - c->mn[0];
        int y0 = (int)ceilf((ty0 - bw) / h) - c->mn[1];
        int y1 = (int)floorf((ty1 + bw) / h) - c->mn[1];
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (x1 > dx - 1) x1 = dx - 1;
        if (y1 > dy - 1) y1 = dy - 1;
        if (x1 < x0 || y1 < y0) continue;
        for (int64_t z = zb; z < ze; ++z) {
            float wz = (float)(c->mn[2] + (int)z) * h;
            if (wz < tz0 - bw || wz > tz1 + bw) continue;
            float bdz = (wz < tz0) ? (tz0 - wz) : ((wz > tz1) ? (wz - tz1) : 0.0f);
            for (int y = y0; y <= y1; ++y) {
                float py = (float)(c->mn[1] + y) * h;
                float bdy = (py < ty0) ? (ty0 - py) : ((py > ty1) ? (py - ty1) : 0.0f);
                size_t row = (size_t)z * zstride + (size_t)y * (size_t)dx;
                for (int x = x0; x <= x1; ++x) {
                    float px = (float)(c->mn[0] + x) * h;
                    float bdx = (px < tx0) ? (tx0 - px) : ((px > tx1) ? (px - tx1) : 0.0f);
                    if (bdx * bdx + bdy * bdy + bdz * bdz > band_sq) continue;
                    float p[3] = {px, py, wz};
                    float d2 = orc_point_triangle_distance_sq(p, tr->v0, tr->v1, tr->v2);
                    if (d2 <= band_sq) {
                        float s = sqrtf(d2);
                        if (s < c->d[row + x]) c->d[row + x] = s;
                    }
                }
            }
        }
    }
}

void orc_splat_band(const orc_tri* t, int64_t nt, float h, float bw,
                    const int32_t min_c[3], const int32_t dims[3], float* d) {
    splat_ctx c = {t, nt, h, bw, min_c, dims, d};
    orc_parallel_for(dims[2], 1, splat_range, &c);
}

/* Binary64 reference of the band: for every voxel, the minimum over all
 * triangles of the binary64 point-triangle distance from its binary32 lattice
 * point (float)(min + i) * h to the binary32 vertices, wherever that minimum
 * is <= bw; 1e30 elsewhere.  Not part of the reference algorithm: the
 * yardstick of the precision study. */
typedef struct {
    const orc_tri* t;
    int64_t nt;
    float h;
    double bw;
    const int32_t* mn;
    const int32_t* dm;
    double* d;
} band64_ctx;

static void band64_range(void* vctx, int64_t zb, int64_t ze, int tid) {
    (void)tid;
    band64_ctx* c = (band64_ctx*)vctx;
    const float h = c->h;
    const double bw = c->bw * (1.0 + 1e-6), band_sq = bw * bw;
    const int dx = c->dm[0], dy = c->dm[1];
    const size_t zstride = (size_t)dy * (size_t)dx;
    for (int64_t ti = 0; ti < c->nt; ++ti) {
        const orc_tri* tr = &c->t[ti];
        double a[3], b[3], e[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = tr->v0[k];
            b[k] = tr->v1[k];
            e[k] = tr->v2[k];
        }
        double lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(a[k], fmin(b[k], e[k])) - bw;
            hi[k] = fmax(a[k], fmax(b[k], e[k])) + bw;
        }
        int x0 = (int)floor(lo[0] / h) - c->mn[0] - 1, x1 = (int)ceil(hi[0] / h) - c->mn[0] + 1;
        int y0 = (int)floor(lo[1] / h) - c->mn[1] - 1, y1 = (int)ceil(hi[1] / h) - c->mn[1] + 1;
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (x1 > dx - 1) x1 = dx - 1;
        if (y1 > dy - 1) y1 = dy - 1;
        for (int64_t z = zb; z < ze; ++z) {
            const double pz = (double)((float)(c->mn[2] + (int)z) * h);
            if (pz < lo[2] || pz > hi[2]) continue;
            for (int y = y0; y <= y1; ++y) {
                const double py = (double)((float)(c->mn[1] + y) * h);
                if (py < lo[1] || py > hi[1]) continue;
                size_t row = (size_t)z * zstride + (size_t)y * (size_t)dx;
                for (int x = x0; x <= x1; ++x) {
                    const double px = (double)((float)(c->mn[0] + x) * h);
                    if (px < lo[0] || px > hi[0]) continue;
                    const double p[3] = {px, py, pz};
                    const double d2 = orc_point_triangle_distance_sq_f64(p, a, b, e);
                    if (d2 <= band_sq) {
                        const double sd = sqrt(d2);
                        if (sd < c->d[row + x]) c->d[row + x] = sd;
                    }
                }
            }
        }
    }
}

void orc_band_f64(const orc_tri* t, int64_t nt, float h, double bw, const int32_t min_c[3],
                  const int32_t dims[3], double* d) {
    const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    for (int64_t i = 0; i < n; ++i) d[i] = 1e30;
    band64_ctx c = {t, nt, h, bw, min_c, dims, d};
    orc_parallel_for(dims[2], 1, band64_range, &c);
}

/* Supercell coverage of one triangle; cuda_rasterizer.cu:836-859. */
static void csr_cover(const orc_tri* tr, float h, int sup, const int32_t min_c[3],
                      int sdx, int sdy, int* sx0, int* sx1, int* sy0, int* sy1,
                      float xyab[4]) {
    float minx = tri_min(tr->v0[0], tr->v1[0], tr->v2[0]);
    float maxx = tri_max(tr->v0[0], tr->v1[0], tr->v2[0]);
    float miny = tri_min(tr->v0[1], tr->v1[1], tr->v2[1]);
    float maxy = tri_max(tr->v0[1], tr->v1[1], tr->v2[1]);
    xyab[0] = minx;
    xyab[1] = miny;
    xyab[2] = maxx;
    xyab[3] = maxy;
    int cx0 = (int)floorf(minx / h) - min_c[0] - 1;
    int cx1 = (int)ceilf(maxx / h) - min_c[0] + 1;
    int cy0 = (int)floorf(miny / h) - min_c[1] - 1;
    int cy1 = (int)ceilf(maxy / h) - min_c[1] + 1;
    /* C integer division truncates toward zero, as in the reference. */
    int a = cx0 / sup, b = cx1 / sup, c = cy0 / sup, e = cy1 / sup;
    *sx0 = a > 0 ? a : 0;
    *sx1 = b < sdx - 1 ? b : sdx - 1;
    *sy0 = c > 0 ? c : 0;
    *sy1 = e < sdy - 1 ? e : sdy - 1;
}

int64_t orc_csr_entries(const orc_tri* t, int64_t nt, float h, int sup,
                        const int32_t min_c[3], const int32_t dims[3]) {
    int sdx = (dims[0] + sup - 1) / sup, sdy = (dims[1] + sup - 1) / sup;
    int64_t total = 0;
    for (int64_t i = 0; i < nt; ++i) {
        int sx0, sx1, sy0, sy1;
        float ab[4];
        csr_cover(&t[i], h, sup, min_c, sdx, sdy, &sx0, &sx1, &sy0, &sy1, ab);
        if (sx1 < sx0 || sy1 < sy0) continue;
        total += (int64_t)(sx1 - sx0 + 1) * (sy1 - sy0 + 1);
    }
    return total;
}

typedef struct {
    const orc_tri* t;
    const float* xyab;
    const int64_t* start;
    const int32_t* ids;
    int sdx, sup;
    float h, jx, jy;
    int max_hits;
    const int32_t* mn;
    const int32_t* dm;
    uint8_t* sign;
    int32_t* hits_out;
    atomic_int flag;
} parity_ctx;

static int cmp_float(const void* a, const void* b) {
    float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}

/* gpu_parity_sign_columns_kernel; cuda_rasterizer.cu:279-368. */
static void parity_range(void* vctx, int64_t cb, int64_t ce, int tid) {
    (void)tid;
    parity_ctx* c = (parity_ctx*)vctx;
    const int dx = c->dm[0], dy = c->dm[1], dz = c->dm[2];
    const size_t zstride = (size_t)dy * (size_t)dx;
    float* hit = (float*)malloc(sizeof(float) * (size_t)(c->max_hits > 0 ? c->max_hits : 1));
    for (int64_t col = cb; col < ce; ++col) {
        int cx = (int)(col % dx), cy = (int)(col / dx);
        float rx = ((float)(c->mn[0] + cx) + c->jx) * c->h;
        float ry = ((float)(c->mn[1] + cy) + c->jy) * c->h;
        int n = 0, bad = 0, all_hits = 0;
        int sup = (cy / c->sup) * c->sdx + (cx / c->sup);
        for (int64_t e = c->start[sup]; e < c->start[sup + 1]; ++e) {
            int ti = c->ids[e];
            const float* ab = c->xyab + 4 * (size_t)ti;
            if (rx < ab[0] || rx > ab[2] || ry < ab[1] || ry > ab[3]) continue;
            const orc_tri* tr = &c->t[ti];
            float d0 = (tr->v1[0] - tr->v0[0]) * (ry - tr->v0[1]) - (tr->v1[1] - tr->v0[1]) * (rx - tr->v0[0]);
            float d1 = (tr->v2[0] - tr->v1[0]) * (ry - tr->v1[1]) - (tr->v2[1] - tr->v1[1]) * (rx - tr->v1[0]);
            float d2 = (tr->v0[0] - tr->v2[0]) * (ry - tr->v2[1]) - (tr->v0[1] - tr->v2[1]) * (rx - tr->v2[0]);
            int pos = (d0 > 0.0f && d1 > 0.0f && d2 > 0.0f);
            int neg = (d0 < 0.0f && d1 < 0.0f && d2 < 0.0f);
            if (!pos && !neg) continue;
            float area2 = d0 + d1 + d2;
            if (fabsf(area2) < 1e-12f) continue;
            float zh = (d1 * tr->v0[2] + d2 * tr->v1[2] + d0 * tr->v2[2]) / area2;
            ++all_hits;
            if (n < c->max_hits) {
                hit[n++] = zh;
            } else {
                bad = 1;
                break;
            }
        }
        if (c->hits_out) c->hits_out[col] = all_hits;
        size_t base = (size_t)cy * (size_t)dx + (size_t)cx;
        if (bad || (n & 1)) {
            for (int z = 0; z < dz; ++z) c->sign[(size_t)z * zstride + base] = 2;
            atomic_store(&c->flag, 1);
            continue;
        }
        qsort(hit, (size_t)n, sizeof(float), cmp_float);
        int hi = 0, inside = 0;
        for (int z = 0; z < dz; ++z) {
            float zc = (float)(c->mn[2] + z) * c->h;
            while (hi < n && hit[hi] < zc) {
                inside = !inside;
                ++hi;
            }
            c->sign[(size_t)z * zstride + base] = inside ? 1 : 0;
        }
    }
    free(hit);
}

int orc_parity_signs(const orc_tri* t, int64_t nt, float h, float jx, float jy,
                     int max_hits, const int32_t min_c[3], const int32_t dims[3],
                     uint8_t* sign, int32_t* hits_out) {
    const int sup = 4;
    int sdx = (dims[0] + sup - 1) / sup, sdy = (dims[1] + sup - 1) / sup;
    size_t ns = (size_t)sdx * (size_t)sdy;
    int64_t* start = (int64_t*)calloc(ns + 1, sizeof(int64_t));
    float* xyab = (float*)malloc(sizeof(float) * 4 * (size_t)(nt > 0 ? nt : 1));
    int* cov = (int*)malloc(sizeof(int) * 4 * (size_t)(nt > 0 ? nt : 1));
    for (int64_t i = 0; i < nt; ++i) {
        int* cv = cov + 4 * i;
        csr_cover(&t[i], h, sup, min_c, sdx, sdy, &cv[0], &cv[1], &cv[2], &cv[3], xyab + 4 * i);
        if (cv[1] < cv[0] || cv[3] < cv[2]) continue;
        for (int sy = cv[2]; sy <= cv[3]; ++sy)
            for (int sx = cv[0]; sx <= cv[1]; ++sx) ++start[(size_t)sy * sdx + sx + 1];
    }
    for (size_t i = 1; i <= ns; ++i) start[i] += start[i - 1];
    int32_t* ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)(start[ns] > 0 ? start[ns] : 1));
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (ns > 0 ? ns : 1));
    memcpy(cur, start, sizeof(int64_t) * ns);
    for (int64_t i = 0; i < nt; ++i) {
        int* cv = cov + 4 * i;
        if (cv[1] < cv[0] || cv[3] < cv[2]) continue;
        for (int sy = cv[2]; sy <= cv[3]; ++sy)
            for (int sx = cv[0]; sx <= cv[1]; ++sx) ids[cur[(size_t)sy * sdx + sx]++] = (int32_t)i;
    }
    parity_ctx c;
    c.t = t;
    c.xyab = xyab;
    c.start = start;
    c.ids = ids;
    c.sdx = sdx;
    c.sup = sup;
    c.h = h;
    c.jx = jx;
    c.jy = jy;
    c.max_hits = max_hits;
    c.mn = min_c;
    c.dm = dims;
    c.sign = sign;
    c.hits_out = hits_out;
    atomic_init(&c.flag, 0);
    orc_parallel_for((int64_t)dims[0] * dims[1], 256, parity_range, &c);
    free(start);
    free(xyab);
    free(cov);
    free(ids);
    free(cur);
    return atomic_load(&c.flag);
}

/* Exact winding at a lattice voxel.  mode 0 restates the plain FP32 sum of
 * gpu_exact_winding_fixup_kernel (cuda_rasterizer.cu:385-397); modes 1 and 2
 * are the compensated-FP32 and binary64 variants used as precision oracles. */
double orc_winding_at_voxel(const orc_tri* t, int64_t nt, float h,
                            const int32_t min_c[3], const int32_t dims[3],
                            int64_t v, int mode) {
    int x = (int)(v % dims[0]);
    int y = (int)((v / dims[0]) % dims[1]);
    int z = (int)(v / ((int64_t)dims[0] * dims[1]));
    float p[3] = {(float)(min_c[0] + x) * h, (float)(min_c[1] + y) * h,
                  (float)(min_c[2] + z) * h};
    if (mode == 0) {
        float s = 0.0f;
        for (int64_t i = 0; i < nt; ++i) s += orc_triangle_solid_angle(p, t[i].v0, t[i].v1, t[i].v2);
        return (double)(s * ORC_INV_4PI_F);
    }
    if (mode == 1 || mode == 3) {
        float s = 0.0f, comp = 0.0f;
        for (int64_t i = 0; i < nt; ++i) {
            float term = orc_triangle_solid_angle(p, t[i].v0, t[i].v1, t[i].v2);
            float yk = term - comp;
            float tk = s + yk;
            comp = (tk - s) - yk;
            s = tk;
        }
        float w = (s - comp) * ORC_INV_4PI_F;
        if (mode == 1 || fabsf(fabsf(w) - 0.5f) >= 1e-3f) return (double)w;
    }
    double pd[3] = {p[0], p[1], p[2]};
    double s = 0.0;
    for (int64_t i = 0; i < nt; ++i) {
        double a[3] = {t[i].v0[0], t[i].v0[1], t[i].v0[2]};
        double b[3] = {t[i].v1[0], t[i].v1[1], t[i].v1[2]};
        double c[3] = {t[i].v2[0], t[i].v2[1], t[i].v2[2]};
        s += orc_triangle_solid_angle_f64(pd, a, b, c);
    }
    return s / (4.0 * ORC_PI);
}

typedef struct {
    const orc_tri* t;
    int64_t nt;
    float h;
    const int32_t* mn;
    const int32_t* dm;
    uint8_t* sign;
    atomic_llong count;
    int mode;
} fixup_ctx;

static void fixup_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    fixup_ctx* c = (fixup_ctx*)vctx;
    int64_t k = 0;
    for (int64_t i = b; i < e; ++i) {
        if (c->sign[i] != 2) continue;
        double w = orc_winding_at_voxel(c->t, c->nt, c->h, c->mn, c->dm, i, c->mode);
        c->sign[i] = (fabs(w) >= 0.5) ? 1 : 0;
        ++k;
    }
    atomic_fetch_add(&c->count, k);
}

typedef struct {
    const orc_tri* t;
    int64_t nt;
    float h;
    const int32_t* mn;
    const int32_t* dm;
    const int64_t* list;
    double* w;
    int mode;
} wlist_ctx;

static void wlist_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    wlist_ctx* c = (wlist_ctx*)vctx;
    for (int64_t i = b; i < e; ++i)
        c->w[i] = orc_winding_at_voxel(c->t, c->nt, c->h, c->mn, c->dm, c->list[i], c->mode);
}

void orc_winding_voxels(const orc_tri* t, int64_t nt, float h, const int32_t min_c[3],
                        const int32_t dims[3], const int64_t* list, int64_t n, int mode, double* w) {
    wlist_ctx c = {t, nt, h, min_c, dims, list, w, mode};
    orc_parallel_for(n, 16, wlist_range, &c);
}

/* gpu_exact_winding_fixup_kernel; cuda_rasterizer.cu:374-400. */
int64_t orc_winding_fixup(const orc_tri* t, int64_t nt, float h,
                          const int32_t min_c[3], const int32_t dims[3], uint8_t* sign, int mode) {
    fixup_ctx c;
    c.t = t;
    c.nt = nt;
    c.h = h;
    c.mn = min_c;
    c.dm = dims;
    c.sign = sign;
    c.mode = mode;
    atomic_init(&c.count, 0);
    int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    orc_parallel_for(n, 4096, fixup_range, &c);
    return atomic_load(&c.count);
}

/* gpu_eikonal_fsm_sweep_kernel; cuda_rasterizer.cu:433-513.
 *
 * Schedule.  The reference walks each line sequentially (Gauss-Seidel along
 * the swept axis) while other lines run concurrently and read each other's
 * cells with no ordering.  Its Godunov update is not monotone in the largest
 * neighbour (the 3-D discriminant halves the squared differences, :416-418),
 * so the result depends on which cross-line values a read happens to see.
 * This restatement fixes one legitimate interleaving of that kernel: every
 * cross-line read sees the value from before the sweep (src), the along-line
 * neighbour behind the walker sees the value just written.  dst receives the
 * whole swept field.  The HIP sweeps implement the same schedule, so the two
 * agree bit-for-bit after any number of sweeps. */
typedef struct {
    const float* src;
    float* dst;
    int dx, dy, dz, axis, dir, ntx;
    float h;
    atomic_int changed;
} sweepdb_ctx;

#define ORC_SWEEP_TILE 64

/* Lines are independent under this schedule (they read only src and write
 * their own cells), so they are split over threads.  An x-sweep walks one
 * line per group; y- and z-sweeps walk ORC_SWEEP_TILE adjacent lines of one
 * x-row together, step by step, so every step reads contiguous rows.  The
 * arithmetic of each cell is the same in every arrangement. */
static void sweepdb_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    sweepdb_ctx* c = (sweepdb_ctx*)vctx;
    const float S = 1e30f;
    const int dx = c->dx, dy = c->dy, dz = c->dz, axis = c->axis, dir = c->dir;
    const size_t ys = (size_t)dx, zs = (size_t)dx * (size_t)dy;
    const float h = c->h;
    const float* src = c->src;
    float* dst = c->dst;
    int changed = 0;
    float behind[ORC_SWEEP_TILE];
    for (int64_t g = b; g < e; ++g) {
        if (axis == 0) {
            const int y = (int)(g % dy), z = (int)(g / dy);
            const size_t row = (size_t)z * zs + (size_t)y * ys;
            float bh = S; /* new value of the previous cell along the walk */
            for (int s = 0; s < dx; ++s) {
                const int x = dir > 0 ? s : dx - 1 - s;
                const size_t i = row + (size_t)x;
                const float cur = src[i];
                const float lo = x > 0 ? (dir > 0 ? bh : src[i - 1]) : S;
                const float hi = x < dx - 1 ? (dir > 0 ? src[i + 1] : bh) : S;
                const float ax = fminf(lo, hi);
                const float ay = fminf(y > 0 ? src[i - ys] : S, y < dy - 1 ? src[i + ys] : S);
                const float az = fminf(z > 0 ? src[i - zs] : S, z < dz - 1 ? src[i + zs] : S);
                const float u = orc_godunov(ax, ay, az, h);
                float nv = cur;
                if (u < cur) {
                    nv = u;
                    changed = 1;
                }
                dst[i] = nv;
                bh = nv;
            }
            continue;
        }
        const int xt = (int)(g % c->ntx), o = (int)(g / c->ntx);
        const int x0 = xt * ORC_SWEEP_TILE;
        const int x1 = x0 + ORC_SWEEP_TILE < dx ? x0 + ORC_SWEEP_TILE : dx;
        const int len = axis == 1 ? dy : dz;
        const size_t as = axis == 1 ? ys : zs;
        int y = axis == 2 ? o : 0, z = axis == 1 ? o : 0;
        for (int k = 0; k < x1 - x0; ++k) behind[k] = S;
        for (int s = 0; s < len; ++s) {
            const int q = dir > 0 ? s : len - 1 - s;
            if (axis == 1) y = q;
            else z = q;
            const size_t row = (size_t)z * zs + (size_t)y * ys;
            const int lo_ok = q > 0, hi_ok = q < len - 1;
            for (int x = x0; x < x1; ++x) {
                const size_t i = row + (size_t)x;
                const float cur = src[i];
                const float lo = lo_ok ? (dir > 0 ? behind[x - x0] : src[i - as]) : S;
                const float hi = hi_ok ? (dir > 0 ? src[i + as] : behind[x - x0]) : S;
                const float ax = fminf(x > 0 ? src[i - 1] : S, x < dx - 1 ? src[i + 1] : S);
                float ay, az;
                if (axis == 1) {
                    ay = fminf(lo, hi);
                    az = fminf(z > 0 ? src[i - zs] : S, z < dz - 1 ? src[i + zs] : S);
                } else {
                    ay = fminf(y > 0 ? src[i - ys] : S, y < dy - 1 ? src[i + ys] : S);
                    az = fminf(lo, hi);
                }
                const float u = orc_godunov(ax, ay, az, h);
                float nv = cur;
                if (u < cur) {
                    nv = u;
                    changed = 1;
                }
                dst[i] = nv;
                behind[x - x0] = nv;
            }
        }
    }
    if (changed) atomic_store(&c->changed, 1);
}

int orc_fsm_sweep_db(const float* src, float* dst, const int32_t dims[3], float h,
                     int axis, int dir) {
    sweepdb_ctx c;
    c.src = src;
    c.dst = dst;
    c.dx = dims[0];
    c.dy = dims[1];
    c.dz = dims[2];
    c.axis = axis;
    c.dir = dir;
    c.h = h;
    c.ntx = (dims[0] + ORC_SWEEP_TILE - 1) / ORC_SWEEP_TILE;
    atomic_init(&c.changed, 0);
    int64_t groups;
    int64_t chunk = 1;
    if (axis == 0) {
        groups = (int64_t)dims[1] * dims[2];
        chunk = 64;
    } else {
        groups = (int64_t)c.ntx * (axis == 1 ? dims[2] : dims[1]);
    }
    orc_parallel_for(groups, chunk, sweepdb_range, &c);
    return atomic_load(&c.changed);
}

/* In-place convenience wrapper of the same schedule. */
int orc_fsm_sweep(float* g, const int32_t dims[3], float h, int axis, int dir) {
    size_t n = (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2];
    float* tmp = (float*)malloc(sizeof(float) * n);
    int ch = orc_fsm_sweep_db(g, tmp, dims, h, axis, dir);
    memcpy(g, tmp, sizeof(float) * n);
    free(tmp);
    return ch;
}

/* Plain in-place line-after-line sweep (each line sees the lines walked
 * before it); kept to show the schedule dependence in the tests. */
int orc_fsm_sweep_inplace(float* g, const int32_t dims[3], float h, int axis, int dir) {
    const float S = 1e30f;
    const int dx = dims[0], dy = dims[1], dz = dims[2];
    const size_t xs = 1, ys = (size_t)dx, zs = (size_t)dx * (size_t)dy;
    const int len = axis == 0 ? dx : (axis == 1 ? dy : dz);
    const int64_t nl = axis == 0 ? (int64_t)dy * dz : (axis == 1 ? (int64_t)dx * dz : (int64_t)dx * dy);
    int changed = 0;
    for (int64_t line = 0; line < nl; ++line) {
        int x = 0, y = 0, z = 0;
        if (axis == 0) {
            y = (int)(line % dy);
            z = (int)(line / dy);
        } else if (axis == 1) {
            x = (int)(line % dx);
            z = (int)(line / dx);
        } else {
            x = (int)(line % dx);
            y = (int)(line / dx);
        }
        for (int s = 0; s < len; ++s) {
            int q = dir > 0 ? s : len - 1 - s;
            if (axis == 0) x = q;
            else if (axis == 1) y = q;
            else z = q;
            size_t i = (size_t)z * zs + (size_t)y * ys + (size_t)x * xs;
            float cur = g[i];
            float ax = fminf(x > 0 ? g[i - xs] : S, x < dx - 1 ? g[i + xs] : S);
            float ay = fminf(y > 0 ? g[i - ys] : S, y < dy - 1 ? g[i + ys] : S);
            float az = fminf(z > 0 ? g[i - zs] : S, z < dz - 1 ? g[i + zs] : S);
            float u = orc_godunov(ax, ay, az, h);
            if (u < cur) {
                g[i] = u;
                changed = 1;
            }
        }
    }
    return changed;
}

/* Classical fast sweeping with the 8 alternating orderings (Zhao 2005), using
 * the same Godunov update; the second fixed-point oracle for the sweeps. */
int orc_fsm_sweep_octant(float* g, const int32_t dims[3], float h, int sx, int sy, int sz) {
    const float S = 1e30f;
    const int dx = dims[0], dy = dims[1], dz = dims[2];
    const size_t ys = (size_t)dx, zs = (size_t)dx * (size_t)dy;
    int changed = 0;
    for (int kz = 0; kz < dz; ++kz) {
        int z = sz > 0 ? kz : dz - 1 - kz;
        for (int ky = 0; ky < dy; ++ky) {
            int y = sy > 0 ? ky : dy - 1 - ky;
            for (int kx = 0; kx < dx; ++kx) {
                int x = sx > 0 ? kx : dx - 1 - kx;
                size_t i = (size_t)z * zs + (size_t)y * ys + (size_t)x;
                float cur = g[i];
                float ax = fminf(x > 0 ? g[i - 1] : S, x < dx - 1 ? g[i + 1] : S);
                float ay = fminf(y > 0 ? g[i - ys] : S, y < dy - 1 ? g[i + ys] : S);
                float az = fminf(z > 0 ? g[i - zs] : S, z < dz - 1 ? g[i + zs] : S);
                float u = orc_godunov(ax, ay, az, h);
                if (u < cur) {
                    g[i] = u;
                    changed = 1;
                }
            }
        }
    }
    return changed;
}

int orc_fsm_octant_round(float* d, const int32_t dims[3], float h) {
    int ch = 0;
    for (int sz = 1; sz >= -1; sz -= 2)
        for (int sy = 1; sy >= -1; sy -= 2)
            for (int sx = 1; sx >= -1; sx -= 2) ch |= orc_fsm_sweep_octant(d, dims, h, sx, sy, sz);
    return ch;
}

/* gpu_count_far_sentinels_kernel; cuda_rasterizer.cu:523-534. */
static void count_far_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    elem_ctx* c = (elem_ctx*)vctx;
    int64_t k = 0;
    for (int64_t i = b; i < e; ++i) k += (c->d[i] >= 1e29f);
    atomic_fetch_add(&c->count, k);
}

int64_t orc_count_far(const float* d, int64_t n) {
    elem_ctx c;
    c.d = (float*)d;
    atomic_init(&c.count, 0);
    orc_parallel_for(n, 1 << 20, count_far_range, &c);
    return atomic_load(&c.count);
}

/* gpu_clamp_far_sentinels_kernel; cuda_rasterizer.cu:544-556.  intended=1
 * stores the unsigned magnitude so [F] yields -bw inside (the behaviour the
 * reference comment at :536-543 describes); intended=0 keeps the reference's
 * pre-signed value, which [F] (:659) negates again. */
void orc_clamp_far(float* d, const uint8_t* sign, int64_t n, float bw, int intended) {
    for (int64_t i = 0; i < n; ++i)
        if (d[i] >= 1e29f) d[i] = intended ? bw : ((sign[i] == 1) ? -bw : bw);
}

/* Stratified probe positions of gpu_verify_sign_samples_kernel;
 * cuda_rasterizer.cu:582-586. */
void orc_verify_indices(int64_t total, int samples, int gen, int64_t* out) {
    double off = (gen == 0) ? 0.5 : (gen == 1) ? (1.0 / 6.0) : (5.0 / 6.0);
    for (int i = 0; i < samples; ++i) {
        size_t v = (size_t)(((double)i + off) * (double)total / (double)samples);
        if (v >= (size_t)total) v = (size_t)total - 1;
        out[i] = (int64_t)v;
    }
}

/* gpu_mark_patch_unknown_kernel; cuda_rasterizer.cu:615-646. */
void orc_mark_patch(uint8_t* sign, const int64_t* sites, int n_sites, int r,
                    const int32_t dims[3]) {
    const int dx = dims[0], dy = dims[1], dz = dims[2];
    for (int s = 0; s < n_sites; ++s) {
        int64_t v = sites[s];
        int cx = (int)(v % dx), cy = (int)((v / dx) % dy), cz = (int)(v / ((int64_t)dx * dy));
        int x0 = cx - r < 0 ? 0 : cx - r, y0 = cy - r < 0 ? 0 : cy - r, z0 = cz - r < 0 ? 0 : cz - r;
        int x1 = cx + r < dx - 1 ? cx + r : dx - 1;
        int y1 = cy + r < dy - 1 ? cy + r : dy - 1;
        int z1 = cz + r < dz - 1 ? cz + r : dz - 1;
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x)
                    sign[(size_t)z * dx * dy + (size_t)y * dx + x] = 2;
    }
}

/* gpu_apply_signs_kernel; cuda_rasterizer.cu:650-661. */
static void apply_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    elem_ctx* c = (elem_ctx*)vctx;
    for (int64_t i = b; i < e; ++i) {
        float v = c->d[i];
        if (v > 1e29f) v = 1e6f;
        c->d[i] = (c->sign[i] == 1) ? -v : v;
    }
}

void orc_apply_signs(float* d, const uint8_t* sign, int64_t n) {
    elem_ctx c;
    c.d = d;
    c.sign = sign;
    orc_parallel_for(n, 1 << 20, apply_range, &c);
}

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

/* FSM driver: reference round structure (:1010-1063) with optional
 * run-to-convergence. */
/* One round x+, x-, y+, y-, z+, z- through the second buffer tmp: six
 * sweeps, so the field is back in d afterwards. */
static int fsm_round_db(float* d, float* tmp, const int32_t dims[3], float h) {
    int ch = 0;
    ch |= orc_fsm_sweep_db(d, tmp, dims, h, 0, +1);
    ch |= orc_fsm_sweep_db(tmp, d, dims, h, 0, -1);
    ch |= orc_fsm_sweep_db(d, tmp, dims, h, 1, +1);
    ch |= orc_fsm_sweep_db(tmp, d, dims, h, 1, -1);
    ch |= orc_fsm_sweep_db(d, tmp, dims, h, 2, +1);
    ch |= orc_fsm_sweep_db(tmp, d, dims, h, 2, -1);
    return ch;
}

int orc_fsm_round(float* d, const int32_t dims[3], float h) {
    size_t n = (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2];
    float* tmp = (float*)malloc(sizeof(float) * n);
    int ch = fsm_round_db(d, tmp, dims, h);
    free(tmp);
    return ch;
}

static void fsm_rounds(float* d, const int32_t dims[3], float h, float bw,
                       const uint8_t* sign, const orc_params* p, orc_stats* st) {
    int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    float* tmp = (float*)malloc(sizeof(float) * (size_t)n);
    int rounds = 0;
    const int octant = p->fsm_converge == 2;
    int cap = p->fsm_converge ? p->fsm_round_cap : p->max_rounds;
    for (int r = 0; r < cap; ++r) {
        int ch = octant ? orc_fsm_octant_round(d, dims, h) : fsm_round_db(d, tmp, dims, h);
        rounds = r + 1;
        if (!ch) break;
    }
    int64_t resid = orc_count_far(d, n);
    int extra = 0;
    while (resid > 0 && extra < p->max_extra_rounds) {
        if (octant)
            (void)orc_fsm_octant_round(d, dims, h);
        else
            (void)fsm_round_db(d, tmp, dims, h);
        ++extra;
        resid = orc_count_far(d, n);
    }
    free(tmp);
    st->eikonal_rounds = rounds + extra;
    st->extra_rounds = extra;
    st->safe_clamped = 0;
    if (resid > 0) {
        st->safe_clamped = resid;
        orc_clamp_far(d, sign, n, bw, p->clamp_intended_sign);
        resid = orc_count_far(d, n);
    }
    st->far_residual = resid;
}

/* rhi_compute_mesh_sdf_fast; cuda_rasterizer.cu:796-1159. */
int orc_fast_sdf(const orc_tri* t, int64_t nt, float h, const int32_t min_c[3],
                 const int32_t max_c[3], const orc_params* p, float* out,
                 uint8_t* sign_out, orc_stats* st) {
    memset(st, 0, sizeof(*st));
    int32_t dims[3] = {max_c[0] - min_c[0] + 1, max_c[1] - min_c[1] + 1, max_c[2] - min_c[2] + 1};
    int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0 || nt <= 0 || n == 0) {
        st->status = 1;
        st->decline_reason = 1;
        return 1;
    }
    if (n > 2000000000LL || nt * (int64_t)dims[2] > 2000000000LL ||
        (int64_t)dims[0] * dims[1] > 2000000000LL || (int64_t)dims[0] * dims[2] > 2000000000LL ||
        (int64_t)dims[1] * dims[2] > 2000000000LL) {
        st->status = 1;
        st->decline_reason = 1;
        return 1;
    }
    float bw = (float)p->band_voxels * h;
    int64_t entries = orc_csr_entries(t, nt, h, p->sup, min_c, dims);
    st->csr_entries = entries;
    if (entries <= 0 || entries > p->csr_max_entries) {
        st->status = 1;
        st->decline_reason = 2;
        return 1;
    }
    double t0 = now_ms();
    float* d = (float*)malloc(sizeof(float) * (size_t)n);
    uint8_t* sg = (uint8_t*)malloc((size_t)n);
    orc_fill(d, n, 1e30f);
    orc_splat_band(t, nt, h, bw, min_c, dims, d);
    double t1 = now_ms();
    st->ms_splat = t1 - t0;
    st->unknown_flag = orc_parity_signs(t, nt, h, p->jitter_x, p->jitter_y, p->parity_max_hits,
                                        min_c, dims, sg, NULL);
    double t2 = now_ms();
    st->ms_parity = t2 - t1;
    if (st->unknown_flag) st->fixup_voxels = orc_winding_fixup(t, nt, h, min_c, dims, sg, p->winding_mode);
    double t3 = now_ms();
    st->ms_fixup = t3 - t2;
    fsm_rounds(d, dims, h, bw, sg, p, st);
    double t4 = now_ms();
    st->ms_eikonal = t4 - t3;

    int ok = 0;
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)p->verify_samples);
    int64_t* sites = (int64_t*)malloc(sizeof(int64_t) * (size_t)p->verify_samples);
    double* wv = (double*)malloc(sizeof(double) * (size_t)p->verify_samples);
    for (int gen = 0; gen < p->verify_gens; ++gen) {
        double tv = now_ms();
        orc_verify_indices(n, p->verify_samples, gen, idx);
        /* every sample's winding number (in parallel), then the mismatches in
         * sample order */
        int64_t ns = 0;
        for (int i = 0; i < p->verify_samples; ++i)
            if (sg[idx[i]] != 2) idx[ns++] = idx[i];
        orc_winding_voxels(t, nt, h, min_c, dims, idx, ns, p->winding_mode, wv);
        int m = 0;
        for (int64_t i = 0; i < ns; ++i) {
            int inside = fabs(wv[i]) >= 0.5;
            if ((sg[idx[i]] == 1) != inside) sites[m++] = idx[i];
        }
        st->probes++;
        st->last_probe_m = m;
        st->ms_verify += now_ms() - tv;
        if (m == 0) {
            ok = 1;
            break;
        }
        if (m > p->max_sites_per_probe) break;
        if (gen == p->verify_gens - 1) break;
        double tp = now_ms();
        orc_mark_patch(sg, sites, m, p->patch_radius, dims);
        st->patched_voxels += orc_winding_fixup(t, nt, h, min_c, dims, sg, p->winding_mode);
        st->ms_patch += now_ms() - tp;
        st->sites += m;
        st->patch_rounds++;
    }
    free(idx);
    free(sites);
    free(wv);
    if (!ok) {
        free(d);
        free(sg);
        st->status = 1;
        st->decline_reason = 3;
        st->ms_total = now_ms() - t0;
        return 1;
    }
    double ta = now_ms();
    orc_apply_signs(d, sg, n);
    st->ms_apply = now_ms() - ta;
    st->ms_total = now_ms() - t0;
    memcpy(out, d, sizeof(float) * (size_t)n);
    if (sign_out) memcpy(sign_out, sg, (size_t)n);
    free(d);
    free(sg);
    st->status = 0;
    return 0;
}

typedef struct {
    const orc_tri* t;
    int64_t nt;
    float h;
    const int32_t* mn;
    const int32_t* dm;
    uint8_t* sign;
    int mode;
} exact_ctx;

static void exact_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    exact_ctx* c = (exact_ctx*)vctx;
    for (int64_t i = b; i < e; ++i) {
        double w = orc_winding_at_voxel(c->t, c->nt, c->h, c->mn, c->dm, i, c->mode);
        c->sign[i] = fabs(w) >= 0.5 ? 1 : 0;
    }
}

int orc_plain_exact_sdf(const orc_tri* t, int64_t nt, float h, const int32_t min_c[3],
                        const int32_t max_c[3], int winding_mode, float* out,
                        uint8_t* sign_out, int* rounds_out) {
    int32_t dims[3] = {max_c[0] - min_c[0] + 1, max_c[1] - min_c[1] + 1, max_c[2] - min_c[2] + 1};
    int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    uint8_t* sg = (uint8_t*)malloc((size_t)n);
    exact_ctx c = {t, nt, h, min_c, dims, sg, winding_mode};
    orc_parallel_for(n, 1024, exact_range, &c);
    float bw = 8.0f * h;
    orc_fill(out, n, 1e30f);
    orc_splat_band(t, nt, h, bw, min_c, dims, out);
    orc_params p;
    orc_params_default(&p);
    orc_stats st;
    memset(&st, 0, sizeof(st));
    fsm_rounds(out, dims, h, bw, sg, &p, &st);
    if (rounds_out) *rounds_out = st.eikonal_rounds;
    orc_apply_signs(out, sg, n);
    if (sign_out) memcpy(sign_out, sg, (size_t)n);
    free(sg);
    return 0;
}

/* ------------------------------------------------------------------ */
/* legacy exact kernel                                                  */
/* ------------------------------------------------------------------ */
typedef struct {
    const orc_tri* t;
    int64_t nt;
    float h;
    const int32_t* mn;
    const int32_t* dm;
    const float* tb; /* per tile bmin[3], bmax[3] */
    int ntiles;
    float* out;
} legacy_ctx;

/* mesh_to_sdf_kernel (cuda_rasterizer.cu:97-185): per voxel min distance
 * with the per-128-triangle tile bbox cull (:140-145) and the plain FP32
 * solid-angle sum over every triangle (:164-174). */
static void legacy_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    legacy_ctx* c = (legacy_ctx*)vctx;
    const int dx = c->dm[0], dy = c->dm[1];
    for (int64_t i = b; i < e; ++i) {
        int x = (int)(i % dx), y = (int)((i / dx) % dy), z = (int)(i / ((int64_t)dx * dy));
        float p[3] = {(float)(c->mn[0] + x) * c->h, (float)(c->mn[1] + y) * c->h,
                      (float)(c->mn[2] + z) * c->h};
        float mind = 1e12f, sum = 0.0f;
        for (int tt = 0; tt < c->ntiles; ++tt) {
            const float* bb = c->tb + 6 * tt;
            float bdx = (p[0] < bb[0]) ? (bb[0] - p[0]) : ((p[0] > bb[3]) ? (p[0] - bb[3]) : 0.0f);
            float bdy = (p[1] < bb[1]) ? (bb[1] - p[1]) : ((p[1] > bb[4]) ? (p[1] - bb[4]) : 0.0f);
            float bdz = (p[2] < bb[2]) ? (bb[2] - p[2]) : ((p[2] > bb[5]) ? (p[2] - bb[5]) : 0.0f);
            float box = bdx * bdx + bdy * bdy + bdz * bdz;
            int can = box < mind;
            int64_t s0 = (int64_t)tt * 128, s1 = s0 + 128 < c->nt ? s0 + 128 : c->nt;
            for (int64_t k = s0; k < s1; ++k) {
                const orc_tri* tr = &c->t[k];
                if (can) {
                    float d2 = orc_point_triangle_distance_sq(p, tr->v0, tr->v1, tr->v2);
                    if (d2 < mind) mind = d2;
                }
                sum += orc_triangle_solid_angle(p, tr->v0, tr->v1, tr->v2);
            }
        }
        float w = sum * ORC_INV_4PI_F;
        float sgn = (fabsf(w) >= 0.5f) ? -1.0f : 1.0f;
        c->out[i] = sgn * sqrtf(mind);
    }
}

void orc_legacy_sdf(const orc_tri* t, int64_t nt, float h, const int32_t min_c[3],
                    const int32_t max_c[3], float* out) {
    int32_t dims[3] = {max_c[0] - min_c[0] + 1, max_c[1] - min_c[1] + 1, max_c[2] - min_c[2] + 1};
    int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    int ntiles = (int)((nt + 127) / 128);
    float* tb = (float*)malloc(sizeof(float) * 6 * (size_t)(ntiles > 0 ? ntiles : 1));
    /* tile headers, cuda_rasterizer.cu:1201-1215 (bbox part; the centre/area
       fields feed only the disabled dipole branch). */
    for (int tt = 0; tt < ntiles; ++tt) {
        float* bb = tb + 6 * tt;
        bb[0] = bb[1] = bb[2] = 1e30f;
        bb[3] = bb[4] = bb[5] = -1e30f;
        int64_t s0 = (int64_t)tt * 128, s1 = s0 + 128 < nt ? s0 + 128 : nt;
        for (int64_t k = s0; k < s1; ++k)
            for (int j = 0; j < 3; ++j) {
                float lo = fminf(t[k].v0[j], fminf(t[k].v1[j], t[k].v2[j]));
                float hi = fmaxf(t[k].v0[j], fmaxf(t[k].v1[j], t[k].v2[j]));
                bb[j] = fminf(bb[j], lo);
                bb[3 + j] = fmaxf(bb[3 + j], hi);
            }
    }
    legacy_ctx c = {t, nt, h, min_c, dims, tb, ntiles, out};
    orc_parallel_for(n, 512, legacy_range, &c);
    free(tb);
}

/* ------------------------------------------------------------------ */
/* standalone APIs                                                      */
/* ------------------------------------------------------------------ */
typedef struct {
    const float* v;
    const uint32_t* idx;
    int64_t nt;
    const float* pts;
    float* w;
} wpts_ctx;

/* gpu_compute_winding_numbers_kernel; cuda_mesh_ops.cu:774-823 (guard < 1e-7
 * skips the triangle, plain FP32 sum in index order). */
static void wpts_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    wpts_ctx* c = (wpts_ctx*)vctx;
    for (int64_t i = b; i < e; ++i) {
        float px = c->pts[3 * i], py = c->pts[3 * i + 1], pz = c->pts[3 * i + 2];
        float total = 0.0f;
        for (int64_t t = 0; t < c->nt; ++t) {
            const float* A = c->v + 3 * (size_t)c->idx[3 * t];
            const float* B = c->v + 3 * (size_t)c->idx[3 * t + 1];
            const float* C = c->v + 3 * (size_t)c->idx[3 * t + 2];
            float a[3] = {A[0] - px, A[1] - py, A[2] - pz};
            float b2[3] = {B[0] - px, B[1] - py, B[2] - pz};
            float cc[3] = {C[0] - px, C[1] - py, C[2] - pz};
            float la = sqrtf(dotf(a, a)), lb = sqrtf(dotf(b2, b2)), lc = sqrtf(dotf(cc, cc));
            if (la < 1e-7f || lb < 1e-7f || lc < 1e-7f) continue;
            float det = a[0] * (b2[1] * cc[2] - b2[2] * cc[1]) - a[1] * (b2[0] * cc[2] - b2[2] * cc[0]) +
                        a[2] * (b2[0] * cc[1] - b2[1] * cc[0]);
            float dab = dotf(a, b2), dbc = dotf(b2, cc), dca = dotf(cc, a);
            float den = la * lb * lc + dab * lc + dbc * la + dca * lb;
            total += 2.0f * atan2f(det, den);
        }
        c->w[i] = total * ORC_INV_4PI_F;
    }
}

void orc_winding_points(const float* verts, const uint32_t* idx, int64_t nt,
                        const float* pts, int64_t np, float* out_w) {
    wpts_ctx c = {verts, idx, nt, pts, out_w};
    orc_parallel_for(np, 64, wpts_range, &c);
}

/* gpu_query_point_distances_kernel; cuda_mesh_ops.cu:668-703: sqrt of the
 * smallest squared point-triangle distance over every triangle, starting
 * from 1e18 (a NaN distance never replaces the minimum). */
static void qdist_range(void* vctx, int64_t b, int64_t e, int tid) {
    (void)tid;
    wpts_ctx* c = (wpts_ctx*)vctx;
    for (int64_t i = b; i < e; ++i) {
        const float p[3] = {c->pts[3 * i], c->pts[3 * i + 1], c->pts[3 * i + 2]};
        float best = 1e18f;
        for (int64_t t = 0; t < c->nt; ++t) {
            const float* A = c->v + 3 * (size_t)c->idx[3 * t];
            const float* B = c->v + 3 * (size_t)c->idx[3 * t + 1];
            const float* C = c->v + 3 * (size_t)c->idx[3 * t + 2];
            const float d2 = orc_point_triangle_distance_sq(p, A, B, C);
            if (d2 < best) best = d2;
        }
        c->w[i] = sqrtf(best);
    }
}

void orc_query_distances(const float* verts, const uint32_t* idx, int64_t nt,
                         const float* pts, int64_t np, float* out_d) {
    wpts_ctx c = {verts, idx, nt, pts, out_d};
    orc_parallel_for(np, 64, qdist_range, &c);
}

/* gpu_eikonal_update_kernel red-black iterations; cuda_rasterizer.cu:1536-1629.
 * Same-colour voxels never neighbour each other, so each half-step is
 * order-independent. */
void orc_eikonal_redblack(float* g, const int32_t dims[3], float h, int iters) {
    const int dx = dims[0], dy = dims[1], dz = dims[2];
    const size_t ys = (size_t)dx, zs = (size_t)dx * (size_t)dy;
    for (int it = 0; it < iters; ++it) {
        for (int par = 0; par < 2; ++par) {
            for (int z = 1; z < dz - 1; ++z)
                for (int y = 1; y < dy - 1; ++y)
                    for (int x = 1; x < dx - 1; ++x) {
                        if (((x + y + z) & 1) != par) continue;
                        size_t i = (size_t)z * zs + (size_t)y * ys + (size_t)x;
                        float cur = fabsf(g[i]);
                        float sg = (g[i] >= 0.0f) ? 1.0f : -1.0f;
                        float ax = fminf(fabsf(g[i - 1]), fabsf(g[i + 1]));
                        float ay = fminf(fabsf(g[i - ys]), fabsf(g[i + ys]));
                        float az = fminf(fabsf(g[i - zs]), fabsf(g[i + zs]));
                        float u = orc_godunov(ax, ay, az, h);
                        if (u < cur) g[i] = u * sg;
                    }
        }
    }
}

/* dev_hash_01; math_utils.cuh:219-226. */
/* ------------------------------------------------------------------ */
/* mesh orientation (cuda_mesh_topology.cu:253-454, 653-842)             */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t key;
    int64_t tri;
    int dir;
    int edge;
} orc_edge;

static int cmp_edge(const void* a, const void* b) {
    const orc_edge* x = (const orc_edge*)a;
    const orc_edge* y = (const orc_edge*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->tri != y->tri) return x->tri < y->tri ? -1 : 1;
    return x->edge - y->edge;
}

double orc_signed_volume(const float* v, const uint32_t* idx, int64_t nt) {
    double s = 0.0;
    for (int64_t t = 0; t < nt; ++t) {
        const float* a = v + 3 * (size_t)idx[3 * t];
        const float* b = v + 3 * (size_t)idx[3 * t + 1];
        const float* c = v + 3 * (size_t)idx[3 * t + 2];
        s += (double)a[0] * ((double)b[1] * c[2] - (double)b[2] * c[1]) +
             (double)a[1] * ((double)b[2] * c[0] - (double)b[0] * c[2]) +
             (double)a[2] * ((double)b[0] * c[1] - (double)b[1] * c[0]);
    }
    return s / 6.0;
}

/* Links between the two triangles of every edge shared by exactly two
 * triangles (found by sorting the edges by their undirected key), a
 * breadth-first walk of each component from its lowest-numbered triangle
 * assigning flips relative to it, a link its walk contradicts marking the
 * mesh non-orientable; flips applied when any link was traversed the same
 * way by both triangles; the whole mesh reversed when its signed volume is
 * below -1e-5. */
void orc_orient_mesh(const float* v, int64_t nv, uint32_t* idx, int64_t nt, int* is_orientable,
                     int* had_inconsistent, int* was_inverted, double* signed_volume) {
    (void)nv;
    *is_orientable = 1;
    *had_inconsistent = 0;
    *was_inverted = 0;
    *signed_volume = 0.0;
    if (nt <= 0) return;
    orc_edge* e = (orc_edge*)malloc(sizeof(orc_edge) * 3 * (size_t)nt);
    int64_t ne = 0;
    for (int64_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) {
            uint32_t a = idx[3 * t + k], b = idx[3 * t + (k + 1) % 3];
            if (a == b) continue;
            uint32_t mn = a < b ? a : b, mx = a < b ? b : a;
            e[ne].key = ((uint64_t)mn << 32) | mx;
            e[ne].tri = t;
            e[ne].dir = a < b ? 0 : 1;
            e[ne].edge = k;
            ++ne;
        }
    qsort(e, (size_t)ne, sizeof(orc_edge), cmp_edge);
    int64_t* nb = (int64_t*)malloc(sizeof(int64_t) * 3 * (size_t)nt);
    uint8_t* same = (uint8_t*)calloc(3 * (size_t)nt, 1);
    for (int64_t i = 0; i < 3 * nt; ++i) nb[i] = -1;
    for (int64_t i = 0; i < ne;) {
        int64_t j = i;
        while (j < ne && e[j].key == e[i].key) ++j;
        if (j - i == 2 && e[i].tri != e[i + 1].tri) {
            int sm = e[i].dir == e[i + 1].dir;
            nb[3 * e[i].tri + e[i].edge] = e[i + 1].tri;
            nb[3 * e[i + 1].tri + e[i + 1].edge] = e[i].tri;
            same[3 * e[i].tri + e[i].edge] = (uint8_t)sm;
            same[3 * e[i + 1].tri + e[i + 1].edge] = (uint8_t)sm;
            if (sm) *had_inconsistent = 1;
        }
        i = j;
    }
    free(e);
    int8_t* flip = (int8_t*)malloc((size_t)nt);
    int64_t* queue = (int64_t*)malloc(sizeof(int64_t) * (size_t)nt);
    for (int64_t t = 0; t < nt; ++t) flip[t] = -1;
    for (int64_t r = 0; r < nt; ++r) {
        if (flip[r] >= 0) continue;
        int64_t head = 0, tail = 0;
        flip[r] = 0;
        queue[tail++] = r;
        while (head < tail) {
            int64_t t = queue[head++];
            for (int k = 0; k < 3; ++k) {
                int64_t n = nb[3 * t + k];
                if (n < 0) continue;
                int want = flip[t] ^ same[3 * t + k];
                if (flip[n] < 0) {
                    flip[n] = (int8_t)want;
                    queue[tail++] = n;
                } else if (flip[n] != want) {
                    *is_orientable = 0;
                }
            }
        }
    }
    if (*had_inconsistent)
        for (int64_t t = 0; t < nt; ++t)
            if (flip[t]) {
                uint32_t a = idx[3 * t + 1];
                idx[3 * t + 1] = idx[3 * t + 2];
                idx[3 * t + 2] = a;
            }
    double vol = orc_signed_volume(v, idx, nt);
    if (vol < -1e-5) {
        for (int64_t t = 0; t < nt; ++t) {
            uint32_t a = idx[3 * t + 1];
            idx[3 * t + 1] = idx[3 * t + 2];
            idx[3 * t + 2] = a;
        }
        vol = -vol;
        *was_inverted = 1;
    }
    *signed_volume = vol;
    free(nb);
    free(same);
    free(flip);
    free(queue);
}

float orc_hash_01(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return (float)(x & 0x00FFFFFFu) / 16777216.0f;
}

/* gpu_generate_gwn_test_points_kernel; cuda_verification.cu:209-264. */
void orc_gwn_test_points(int mode, uint32_t seed, float* out) {
    for (int i = 0; i < 10000; ++i) {
        uint32_t s0 = seed + (uint32_t)(i * 9973 + 17);
        float* o = out + 3 * i;
        if (mode == 0 || mode == 2) {
            if (i < 5000) {
                float sx = (mode == 0) ? 17.0f : 25.0f;
                float sz = (mode == 0) ? 11.5f : 17.0f;
                o[0] = -8.5f + sx * orc_hash_01(s0 + 1u);
                o[1] = -8.5f + 17.0f * orc_hash_01(s0 + 2u);
                o[2] = -8.5f + sz * orc_hash_01(s0 + 3u);
            } else {
                int done = 0;
                for (int att = 0; att < 64 && !done; ++att) {
                    float x = -35.0f + 70.0f * orc_hash_01(s0 + (uint32_t)(att * 3 + 1));
                    float y = -35.0f + 70.0f * orc_hash_01(s0 + (uint32_t)(att * 3 + 2));
                    float z = -35.0f + 70.0f * orc_hash_01(s0 + (uint32_t)(att * 3 + 3));
                    int inside_box = (mode == 0)
                        ? (fabsf(x) <= 12.0f && fabsf(y) <= 12.0f && z >= -12.0f && z <= 12.0f)
                        : (x >= -12.0f && x <= 20.0f && fabsf(y) <= 12.0f && fabsf(z) <= 12.0f);
                    if (!inside_box) {
                        o[0] = x;
                        o[1] = y;
                        o[2] = z;
                        done = 1;
                    }
                }
                if (!done) {
                    float f = (mode == 0) ? 25.0f : 28.0f;
                    o[0] = o[1] = o[2] = f;
                }
            }
        } else {
            float r = (i < 5000) ? (13.5f * orc_hash_01(s0 + 1u))
                                 : (16.5f + 18.5f * orc_hash_01(s0 + 1u));
            float u = -1.0f + 2.0f * orc_hash_01(s0 + 2u);
            float th = 6.2831853f * orc_hash_01(s0 + 3u);
            float s = sqrtf(fmaxf(0.0f, 1.0f - u * u));
            o[0] = r * (s * cosf(th));
            o[1] = r * (s * sinf(th));
            o[2] = r * u;
        }
    }
}

/* ---- marching cubes and interior volume ---------------------------------- */
/* CudaRasterizer::extract_iso_surface (cuda_marching_cubes.cu:312-556,
 * 829-962) with a deterministic output order: vertices voxel by voxel (index
 * order), x-, y-, then z-edge; triangles cube by cube, in table order, each
 * emitted as (e0, e2, e1).  The reference assigns both orders with atomic
 * counters, so its output is this one up to a permutation of vertices and
 * triangles.  Border voxels are raised to iso + h first (the seal pass, here
 * applied on read), and a corner is inside when value - iso <= 0. */
#include "../rhizome_v24.15_amd/csrc/mc_table.inc"

static const int kMcCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
/* cube edge -> (owner corner, axis) of the lattice edge it lies on */
static const int kMcEdgeOwner[12][2] = {{0, 0}, {1, 1}, {3, 0}, {0, 1}, {4, 0}, {5, 1},
                                        {7, 0}, {4, 1}, {0, 2}, {1, 2}, {2, 2}, {3, 2}};

int orc_mc_table(int c, int* edges) {
    int n = kMcTriStart[c + 1] - kMcTriStart[c];
    for (int k = 0; k < n; ++k) edges[k] = kMcTriEdges[kMcTriStart[c] + k];
    return n;
}

static float mc_val(const float* g, const int32_t d[3], int x, int y, int z, float seal) {
    float v = g[((size_t)z * (size_t)d[1] + (size_t)y) * (size_t)d[0] + (size_t)x];
    if (x == 0 || x == d[0] - 1 || y == 0 || y == d[1] - 1 || z == 0 || z == d[2] - 1)
        if (v < seal) v = seal;
    return v;
}

static int mc_case(const float* g, const int32_t d[3], int x, int y, int z, float iso, float seal) {
    int c = 0;
    for (int k = 0; k < 8; ++k)
        if (mc_val(g, d, x + kMcCorner[k][0], y + kMcCorner[k][1], z + kMcCorner[k][2], seal) - iso <= 0.0f)
            c |= 1 << k;
    return c;
}

void orc_mc_count(const float* g, const int32_t d[3], float h, float iso, int64_t* nv, int64_t* nt) {
    const float seal = iso + h;
    int64_t v = 0, t = 0;
    for (int z = 0; z < d[2]; ++z)
        for (int y = 0; y < d[1]; ++y)
            for (int x = 0; x < d[0]; ++x) {
                const int s0 = mc_val(g, d, x, y, z, seal) - iso <= 0.0f;
                if (x + 1 < d[0] && s0 != (mc_val(g, d, x + 1, y, z, seal) - iso <= 0.0f)) ++v;
                if (y + 1 < d[1] && s0 != (mc_val(g, d, x, y + 1, z, seal) - iso <= 0.0f)) ++v;
                if (z + 1 < d[2] && s0 != (mc_val(g, d, x, y, z + 1, seal) - iso <= 0.0f)) ++v;
                if (x + 1 < d[0] && y + 1 < d[1] && z + 1 < d[2]) {
                    const int c = mc_case(g, d, x, y, z, iso, seal);
                    t += (kMcTriStart[c + 1] - kMcTriStart[c]) / 3;
                }
            }
    *nv = v;
    *nt = t;
}

/* verts: 3 * nv floats, idx: 3 * nt entries (sizes from orc_mc_count). */
void orc_mc_extract(const float* g, const int32_t d[3], float h, const float origin[3], float iso,
                    float* verts, uint32_t* idx) {
    const float seal = iso + h;
    const size_t n = (size_t)d[0] * (size_t)d[1] * (size_t)d[2];
    int32_t* ev = (int32_t*)malloc(3 * n * sizeof(int32_t));  /* vertex id per (voxel, axis) */
    if (!ev) abort();
    for (size_t i = 0; i < 3 * n; ++i) ev[i] = -1;
    int64_t nv = 0, nt = 0;
    for (int z = 0; z < d[2]; ++z)
        for (int y = 0; y < d[1]; ++y)
            for (int x = 0; x < d[0]; ++x) {
                const size_t i = ((size_t)z * (size_t)d[1] + (size_t)y) * (size_t)d[0] + (size_t)x;
                const float v0 = mc_val(g, d, x, y, z, seal) - iso;
                const int s0 = v0 <= 0.0f;
                const float p0[3] = {origin[0] + (float)x * h, origin[1] + (float)y * h, origin[2] + (float)z * h};
                for (int a = 0; a < 3; ++a) {
                    const int q[3] = {x + (a == 0), y + (a == 1), z + (a == 2)};
                    if (q[a] >= d[a]) continue;
                    const float v1 = mc_val(g, d, q[0], q[1], q[2], seal) - iso;
                    if (s0 == (v1 <= 0.0f)) continue;
                    const float den = v0 - v1;
                    const float t = (fabsf(den) > 1e-7f) ? fminf(fmaxf(v0 / den, 0.001f), 0.999f) : 0.5f;
                    float* o = verts + 3 * nv;
                    o[0] = p0[0];
                    o[1] = p0[1];
                    o[2] = p0[2];
                    o[a] = p0[a] + t * h;
                    ev[3 * i + (size_t)a] = (int32_t)nv++;
                }
            }
    for (int z = 0; z + 1 < d[2]; ++z)
        for (int y = 0; y + 1 < d[1]; ++y)
            for (int x = 0; x + 1 < d[0]; ++x) {
                const int c = mc_case(g, d, x, y, z, iso, seal);
                for (int k = kMcTriStart[c]; k < kMcTriStart[c + 1]; k += 3) {
                    int32_t id[3];
                    for (int j = 0; j < 3; ++j) {
                        const int e = kMcTriEdges[k + j];
                        const int* oc = kMcCorner[kMcEdgeOwner[e][0]];
                        const size_t o = ((size_t)(z + oc[2]) * (size_t)d[1] + (size_t)(y + oc[1])) * (size_t)d[0] +
                                         (size_t)(x + oc[0]);
                        id[j] = ev[3 * o + (size_t)kMcEdgeOwner[e][1]];
                    }
                    if (id[0] < 0 || id[1] < 0 || id[2] < 0 || id[0] == id[1] || id[1] == id[2] || id[2] == id[0])
                        continue;
                    idx[3 * nt] = (uint32_t)id[0];
                    idx[3 * nt + 1] = (uint32_t)id[2];
                    idx[3 * nt + 2] = (uint32_t)id[1];
                    ++nt;
                }
            }
    free(ev);
}

/* gpu_count_interior_voxels_kernel (ui/viewport/cuda_viewport_gl.cu:84-110):
 * voxels with value <= 0 (NaN is not inside). */
int64_t orc_interior_count(const float* g, int64_t n) {
    int64_t c = 0;
    for (int64_t i = 0; i < n; ++i) c += g[i] <= 0.0f;
    return c;
}
