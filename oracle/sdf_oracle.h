/*
 * sdf_oracle.h -- CPU restatement of the reference mesh-to-SDF "Fast SDF"
 * path (golemid/rhizome_v24.15, native/src/engine/sdf/cuda_rasterizer.cu and
 * native/src/core/math_utils.cuh).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never calls it and has no
 * CPU fallback.
 *
 * Arithmetic: plain C, IEEE binary32/binary64, compiled with
 * -ffp-contract=off so every a*b+c rounds twice, exactly as the HIP kernels
 * that are compared bit-for-bit against it (band distances, parity z-hits,
 * Godunov updates).  atan2f is the host libm's; winding sums are compared by
 * their |w| >= 0.5 decision, never bitwise.
 *
 * Pinning (what checks each function; the tests live in tests/test_oracle_*.py):
 *   orc_point_triangle_distance_sq   brute-force FP64 minimisation over the
 *                                    triangle, vertex/edge/face closed forms
 *   orc_triangle_solid_angle(_f64)   octant = pi/2, regular tetrahedron faces
 *                                    sum to 4 pi, closed mesh 4 pi / 0, flip
 *                                    negates
 *   orc_godunov                      closed forms (s,inf,inf)->s+h,
 *                                    (s,s,inf)->s+h/sqrt2, (s,s,s)->s+h/sqrt3
 *   orc_fsm_*                        planar seeds give the exact linear field,
 *                                    idempotent on a converged field, only
 *                                    lowers values, converged far field within
 *                                    a voxel of the Euclidean distance.  The
 *                                    fixed point depends on the sweep schedule
 *                                    (the halved 3D discriminant makes the
 *                                    update non-monotone in s2): 6-axis and
 *                                    8-ordering schedules differ by up to about
 *                                    a voxel, so the schedule itself is pinned
 *                                    (double-buffered lines, Gauss-Seidel
 *                                    along each line)
 *   orc_fsm_sweep_octant / _round    planar seeds give the exact linear field
 *                                    in one round; every ordering equals an
 *                                    independent diagonal-level traversal
 *                                    bit for bit, and a planted level
 *                                    reversal differs
 *   orc_parity_signs                 brute-force ray/triangle counting; equals
 *                                    exact winding sign away from the surface on
 *                                    watertight meshes
 *   orc_winding_points               reference Test 5 fixtures (open box,
 *                                    inverted sphere, two cubes) with the
 *                                    reference hash point generator
 *   orc_query_distances              FP64 brute-force minimum over the mesh,
 *                                    closed forms (height over a face,
 *                                    distance to an edge / vertex)
 *   orc_mc_*                         table: cut edges, closed and consistently
 *                                    oriented in every case; an independent
 *                                    per-cube extraction (same triangles);
 *                                    sphere: closed 2-manifold, genus 0,
 *                                    enclosed volume; plane: exact positions
 *   orc_interior_count               direct count
 *   orc_orient_mesh / signed_volume  closed forms: seeded flipped faces come
 *                                    back exactly, an inside-out sphere is
 *                                    inverted, a Moebius band is
 *                                    non-orientable, a nested inside-out
 *                                    shell keeps its winding, unit cube and
 *                                    tetrahedron volumes; a planted error in
 *                                    the link direction fails them
 *   orc_eikonal_redblack             linear field from a planar seed, cells
 *                                    with no interior neighbours untouched,
 *                                    signs kept, sweep order changes results
 *   orc_verify_indices               closed form (double arithmetic)
 *   orc_patch_box_*                  closed form
 *   orc_splat_band                   equals the brute-force min over all
 *                                    triangles wherever that min <= band
 *   orc_legacy_sdf                   |p|-R on a sphere (Test 12 bound)
 *   oracle/mlp.py (proposer network) central finite differences of the FP64
 *                                    loss, torch's bfloat16 rounding, a planted
 *                                    mask error that must fail
 * The nvcc FMA-contracted bit patterns of the reference build are not
 * reproducible here (no nvcc, no glm): parity unpinned at the ulp level,
 * absorbed by the 1e-3-voxel distance tolerance.
 */
#ifndef RHZ_SDF_ORACLE_H
#define RHZ_SDF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 48-byte triangle, same layout as the reference CudaTriangle
 * (cuda_types.hpp:18-21: v0, v1, v2, normal as 4 x vec3). */
typedef struct orc_tri {
    float v0[3], v1[3], v2[3], n[3];
} orc_tri;

/* Policy constants; orc_params_default() returns the reference values
 * (cuda_rasterizer.cu:74-82, :825, :1010, :1027, :298-299, :867). */
typedef struct orc_params {
    int64_t fast_min_voxels;   /* 32Mi gate, :74 */
    int band_voxels;           /* 8, :75 */
    int parity_max_hits;       /* 128, :76 */
    int verify_samples;        /* 4096, :79 */
    int verify_gens;           /* 3, :80 */
    int max_sites_per_probe;   /* 64, :81 */
    int patch_radius;          /* 24, :82 */
    int max_rounds;            /* 8, :1010 */
    int max_extra_rounds;      /* 16, :1027 */
    int sup;                   /* 4, :825 */
    int64_t csr_max_entries;   /* 1e8, :867 */
    float jitter_x, jitter_y;  /* 0.037, 0.051, :298-299 */
    int clamp_intended_sign;   /* 1: clamp writes +bw, apply negates inside
                                  (intended, -bw inside); 0: reference
                                  double-sign behaviour (:553 + :659) */
    int fsm_converge;          /* 1: run rounds until nothing changes
                                  (cap fsm_round_cap); 0: reference cap;
                                  2: rounds of the 8 classical orderings
                                  (orc_fsm_octant_round) until nothing
                                  changes (cap fsm_round_cap) */
    int fsm_round_cap;
    int winding_mode;          /* sign decisions of fixup/verify: 0 plain FP32
                                  (reference), 3 compensated FP32 with a
                                  binary64 recompute where ||w|-0.5| < 1e-3 */
} orc_params;

typedef struct orc_stats {
    int status;                /* 0 ok, 1 declined (caller falls back) */
    int decline_reason;        /* 0 none, 1 int range, 2 csr, 3 verify */
    int unknown_flag;
    int64_t fixup_voxels;
    int eikonal_rounds;
    int extra_rounds;
    int64_t far_residual;
    int64_t safe_clamped;
    int probes;
    int sites;
    int64_t patched_voxels;
    int patch_rounds;
    int last_probe_m;
    int64_t csr_entries;
    /* wall time of each stage of orc_fast_sdf (ms) */
    double ms_splat;           /* fill + band splat */
    double ms_parity;
    double ms_fixup;
    double ms_eikonal;         /* rounds, audit, extension, clamp */
    double ms_verify;          /* exact winding at the stratified samples */
    double ms_patch;           /* patch boxes and their exact re-evaluation */
    double ms_apply;
    double ms_total;
} orc_stats;

void orc_params_default(orc_params* p);
int orc_num_threads(void);
void orc_set_num_threads(int n);

/* ---- L0 math (math_utils.cuh) ---- */
float orc_point_triangle_distance_sq(const float p[3], const float a[3],
                                     const float b[3], const float c[3]);
double orc_point_triangle_distance_sq_f64(const double p[3], const double a[3],
                                          const double b[3], const double c[3]);
float orc_triangle_solid_angle(const float p[3], const float v0[3],
                               const float v1[3], const float v2[3]);
double orc_triangle_solid_angle_f64(const double p[3], const double v0[3],
                                    const double v1[3], const double v2[3]);
float orc_godunov(float ax, float ay, float az, float h);

/* ---- mesh intake (math_utils.cuh:484-514, cuda_rasterizer.cu:1321-1435,
 *      cuda_mesh_ops.cu:71-255) ---- */
int64_t orc_build_triangles(const float* verts, int64_t nv, const uint32_t* idx,
                            int64_t nt, orc_tri* out);
void orc_mesh_bbox(const float* verts, int64_t nv, float bmin[3], float bmax[3]);
void orc_frame_grid(const float* verts, int64_t nv, float h,
                    int32_t min_c[3], int32_t max_c[3]);

/* ---- stages ([init]..[F]); dims = max-min+1, x fastest ---- */
void orc_fill(float* d, int64_t n, float v);
void orc_band_f64(const orc_tri* t, int64_t nt, float h, double bw, const int32_t min_c[3],
                  const int32_t dims[3], double* d);
void orc_splat_band(const orc_tri* t, int64_t nt, float h, float bw,
                    const int32_t min_c[3], const int32_t dims[3], float* d);
/* Supercell CSR entry count only (decline check, :824-869). */
int64_t orc_csr_entries(const orc_tri* t, int64_t nt, float h, int sup,
                        const int32_t min_c[3], const int32_t dims[3]);
/* Writes 0/1/2 per voxel, returns the unknown flag.  If hits_out is non-NULL
 * it receives the per-column hit count (dims[0]*dims[1] ints). */
int orc_parity_signs(const orc_tri* t, int64_t nt, float h, float jx, float jy,
                     int max_hits, const int32_t min_c[3], const int32_t dims[3],
                     uint8_t* sign, int32_t* hits_out);
/* Exact winding classification of every sign==2 voxel (winding mode as in
 * orc_winding_at_voxel); returns the count. */
int64_t orc_winding_fixup(const orc_tri* t, int64_t nt, float h,
                          const int32_t min_c[3], const int32_t dims[3],
                          uint8_t* sign, int mode);
/* Winding number (sum/(4pi)) at lattice voxel v: mode 0 = reference FP32
 * plain sum, 1 = FP32 Kahan, 2 = FP64, 3 = FP32 Kahan, recomputed in FP64
 * where ||w| - 0.5| < 1e-3. */
double orc_winding_at_voxel(const orc_tri* t, int64_t nt, float h,
                            const int32_t min_c[3], const int32_t dims[3],
                            int64_t v, int mode);
/* orc_winding_at_voxel for a list of voxels (threads over the list). */
void orc_winding_voxels(const orc_tri* t, int64_t nt, float h, const int32_t min_c[3],
                        const int32_t dims[3], const int64_t* list, int64_t n, int mode, double* w);
/* One directional sweep (axis 0/1/2, dir +1/-1) with the pinned schedule:
 * Gauss-Seidel along the swept line, cross-line neighbours from src.
 * Writes every voxel of dst; returns 1 if any voxel decreased. */
int orc_fsm_sweep_db(const float* src, float* dst, const int32_t dims[3], float h,
                     int axis, int dir);
/* One round x+, x-, y+, y-, z+, z- of that schedule, in place. */
int orc_fsm_round(float* d, const int32_t dims[3], float h);
/* Same schedule, in place (uses a temporary copy). */
int orc_fsm_sweep(float* d, const int32_t dims[3], float h, int axis, int dir);
/* Line-after-line in-place variant (a different legitimate interleaving). */
int orc_fsm_sweep_inplace(float* d, const int32_t dims[3], float h, int axis, int dir);
/* Classical 8-ordering sweep (sx,sy,sz in {+1,-1}); returns changed. */
int orc_fsm_sweep_octant(float* d, const int32_t dims[3], float h,
                         int sx, int sy, int sz);
/* The 8 orderings in turn, sz outermost (+1 first), then sy, then sx. */
int orc_fsm_octant_round(float* d, const int32_t dims[3], float h);
int64_t orc_count_far(const float* d, int64_t n);
void orc_clamp_far(float* d, const uint8_t* sign, int64_t n, float bw,
                   int intended_sign);
void orc_verify_indices(int64_t total, int samples, int gen, int64_t* out);
/* Marks the (2r+1)^3 box around each site as unknown (2). */
void orc_mark_patch(uint8_t* sign, const int64_t* sites, int n_sites, int r,
                    const int32_t dims[3]);
void orc_apply_signs(float* d, const uint8_t* sign, int64_t n);

/* ---- whole paths ---- */
/* Reference fast path (rhi_compute_mesh_sdf_fast, :796-1159).  out receives
 * the signed field only on status 0.  sign_out (optional) receives the final
 * sign volume. */
int orc_fast_sdf(const orc_tri* t, int64_t nt, float h, const int32_t min_c[3],
                 const int32_t max_c[3], const orc_params* p, float* out,
                 uint8_t* sign_out, orc_stats* st);
/* "Plain exact" method: exact winding sign for every voxel (mode as in
 * orc_winding_at_voxel), +-band exact splat, FSM to convergence. */
int orc_plain_exact_sdf(const orc_tri* t, int64_t nt, float h,
                        const int32_t min_c[3], const int32_t max_c[3],
                        int winding_mode, float* out, uint8_t* sign_out,
                        int* rounds_out);
/* Legacy exact kernel (mesh_to_sdf_kernel, :97-185, tile headers :1197-1235). */
void orc_legacy_sdf(const orc_tri* t, int64_t nt, float h,
                    const int32_t min_c[3], const int32_t max_c[3], float* out);
/* Standalone winding numbers over an indexed mesh (cuda_mesh_ops.cu:774-823). */
void orc_winding_points(const float* verts, const uint32_t* idx, int64_t nt,
                        const float* pts, int64_t np, float* out_w);
/* Standalone point-to-mesh distances over an indexed mesh (cuda_mesh_ops.cu:668-703). */
void orc_query_distances(const float* verts, const uint32_t* idx, int64_t nt,
                         const float* pts, int64_t np, float* out_d);
/* Marching cubes (cuda_marching_cubes.cu:312-556, 829-962) in a deterministic
 * order: vertices voxel by voxel, x-, y-, z-edge; triangles cube by cube in
 * table order, emitted (e0, e2, e1).  orc_mc_count gives the sizes. */
int orc_mc_table(int c, int* edges);
void orc_mc_count(const float* g, const int32_t dims[3], float h, float iso, int64_t* nv, int64_t* nt);
void orc_mc_extract(const float* g, const int32_t dims[3], float h, const float origin[3], float iso,
                    float* verts, uint32_t* idx);
/* Winding unification and outward orientation, in place
 * (CudaRasterizer::orient_mesh_gpu, cuda_mesh_topology.cu:303-454, 707-842). */
void orc_orient_mesh(const float* v, int64_t nv, uint32_t* idx, int64_t nt, int* is_orientable,
                     int* had_inconsistent, int* was_inverted, double* signed_volume);
/* sum(v0 . (v1 x v2)) / 6 in binary64, triangle order (:653-705). */
double orc_signed_volume(const float* v, const uint32_t* idx, int64_t nt);
/* Voxels with value <= 0 (ui/viewport/cuda_viewport_gl.cu:84-110). */
int64_t orc_interior_count(const float* g, int64_t n);
/* Standalone red-black eikonal (cuda_rasterizer.cu:1536-1629), 8 iterations. */
void orc_eikonal_redblack(float* grid, const int32_t dims[3], float h, int iters);
/* Reference hash (math_utils.cuh:219-226) and Test 5 point generator
 * (cuda_verification.cu:209-264). */
float orc_hash_01(uint32_t x);
void orc_gwn_test_points(int mode, uint32_t seed, float* out_pts /* 10000x3 */);

#ifdef __cplusplus
}
#endif
#endif
