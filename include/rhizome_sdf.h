/*
 * rhizome_sdf.h -- C-ABI of the MI355X (gfx950) mesh-to-SDF library.
 *
 * Plain C types only (pointers, sizes, POD structs); no HIP or torch types.
 * Every pipeline-level entry point replaces one reference interface, cited in
 * its comment (paths relative to the reference repository root).  Status
 * codes: 0 = OK, >0 = declined (the caller falls back, as the reference's
 * fast path does), <0 = error (rhz_last_error() has the message).
 *
 * Threading: calls on one device that share its cached workspace are
 * serialised by a library lock (held for the whole call, host copies
 * included); each call uses the HIP stream it is given (NULL = the legacy
 * default stream).
 */
#ifndef RHIZOME_SDF_H
#define RHIZOME_SDF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RHZ_OK 0
#define RHZ_DECLINED 1
#define RHZ_ERR_ARG (-1)
#define RHZ_ERR_HIP (-2)
#define RHZ_ERR_NOMEM (-3)
#define RHZ_ERR_NODEV (-4)

/* 48-byte triangle; identical layout to the reference CudaTriangle
 * (native/src/engine/sdf/cuda_types.hpp:18-21). */
typedef struct rhz_triangle {
    float v0[3];
    float v1[3];
    float v2[3];
    float normal[3];
} rhz_triangle;

/* Which pipeline rhz_sdf_compute runs. */
enum rhz_method {
    RHZ_METHOD_AUTO = 0,         /* reference gate: FAST if V >= fast_min_voxels, else LEGACY;
                                    a declined FAST falls through to LEGACY
                                    (cuda_rasterizer.cu:1189-1195) */
    RHZ_METHOD_FAST = 1,         /* reference fast path only (may decline) */
    RHZ_METHOD_LEGACY = 2,       /* exact per-voxel distance + winding (cuda_rasterizer.cu:97-185) */
    RHZ_METHOD_PLAIN_EXACT = 3,  /* exact winding sign for every voxel + band + FSM */
    RHZ_METHOD_FAST_LEARNED = 4  /* fast path with risk-ranked exact re-probes */
};

/* Far-field sweep policy. */
enum rhz_fsm_policy {
    RHZ_FSM_REFERENCE = 0, /* <= max_rounds rounds (exit on no change), sentinel audit,
                              <= max_extra_rounds audit-gated rounds, safe clamp
                              (cuda_rasterizer.cu:1010-1063) */
    RHZ_FSM_CONVERGE = 1,  /* rounds until no voxel changes (cap fsm_round_cap) */
    RHZ_FSM_OCTANT = 2     /* rounds of the 8 classical orderings (Zhao 2005,
                              rhz_stage_sweep_octant) until no voxel changes (cap
                              fsm_round_cap); the audit, extension and clamp of
                              RHZ_FSM_REFERENCE follow.  No counterpart in the
                              reference's pipeline (its 6-axis rounds,
                              cuda_rasterizer.cu:1010-1023); the 8-direction
                              form the plain small configuration names */
};

/* Winding-number arithmetic used by every exact-sign decision. */
enum rhz_winding_precision {
    RHZ_WINDING_FP32 = 0,        /* plain FP32 sum (reference arithmetic) */
    RHZ_WINDING_FP32_KAHAN = 1   /* compensated FP32; FP64 recompute where |w-0.5| < fp64_band */
};

/* Constants of the path; rhz_sdf_params_default() returns the reference
 * values (cuda_rasterizer.cu:74-82, :825, :867, :1010, :1027, :298-299). */
typedef struct rhz_sdf_params {
    int32_t struct_size;          /* sizeof(rhz_sdf_params) */
    int32_t method;               /* enum rhz_method */
    int64_t fast_min_voxels;      /* 32Mi */
    int32_t band_voxels;          /* 8 */
    int32_t parity_max_hits;      /* 128 */
    int32_t verify_samples;       /* 4096 */
    int32_t verify_gens;          /* 3 */
    int32_t max_sites_per_probe;  /* 64 */
    int32_t patch_radius;         /* 24 */
    int32_t max_rounds;           /* 8 */
    int32_t max_extra_rounds;     /* 16 */
    int32_t sup;                  /* 4 (supercell edge in columns) */
    int32_t fsm_policy;           /* enum rhz_fsm_policy */
    int64_t csr_max_entries;      /* 1e8 */
    float jitter_x;               /* 0.037 */
    float jitter_y;               /* 0.051 */
    int32_t clamp_intended_sign;  /* 1: clamped voxels ship as -band inside (the behaviour
                                     the reference comment :536-543 states); 0: the
                                     reference's double negation (:553 + :659) */
    int32_t fsm_round_cap;        /* cap for RHZ_FSM_CONVERGE */
    int32_t winding_precision;    /* enum rhz_winding_precision */
    float fp64_band;              /* 1e-3 */
    float reprobe_fraction;       /* FAST_LEARNED: top-ranked fraction of all voxels re-probed
                                     exactly (e.g. 1e-3); < 0 (default): every candidate the
                                     scorer calls positive (logit >= 0, margin >= 0) */
    int32_t learned_max_patch_gens; /* FAST_LEARNED: generations of 26-neighbour growth around
                                     mismatches (0: off; the certificate below makes it redundant) */
    float candidate_radius;       /* FAST_LEARNED: candidates have band distance <= this many
                                     voxels; 0 = the jitter reach |(jitter_x, jitter_y)| (+0.1%),
                                     beyond which a scanline sign cannot be wrong */
    int32_t scorer;               /* FAST_LEARNED: 0 proposer network, 1 planar-margin rule,
                                     2 blind (no ranking: k candidates chosen by a hash of
                                     their voxel index), 3 certificate flag alone (flagged
                                     first, each class in the same hashed order) */
    int32_t learned_certify;      /* FAST_LEARNED: 1 = also probe every candidate whose voxel-to-ray
                                     segment may cross a triangle and that the ranking missed
                                     (0 sign errors on closed meshes); 0 = ranked probes only */
    float far_field_beta;         /* FAST_LEARNED exact probes: triangle clusters farther than
                                     beta x their radius (beta >= 2) enter through a two-term
                                     multipole expansion with a rigorous error bound; voxels the
                                     bound cannot decide go to the brute-force kernel (decisions
                                     unchanged); 0 = brute force for every probe.  Default 4 */
    const float* mlp_weights;     /* FAST_LEARNED: proposer parameters, rhz_mlp_blob_count()
                                     floats (host memory); NULL = the built-in trained set */
} rhz_sdf_params;

/* Counters of one call; the fields of the reference's log line
 * (cuda_rasterizer.cu:1145-1157) plus device stage times. */
typedef struct rhz_sdf_stats {
    int32_t status;
    int32_t decline_reason;       /* 0 none, 1 int range, 2 CSR, 3 verify, 4 memory */
    int32_t method_used;
    int32_t unknown_flag;
    int32_t dims[3];
    int32_t eikonal_rounds;
    int32_t extra_rounds;
    int32_t sweeps;
    int64_t n_triangles;
    int64_t fixup_voxels;
    int64_t far_residual;
    int64_t safe_clamped;
    int32_t probes;
    int32_t sites;
    int64_t patched_voxels;
    int32_t patch_rounds;
    int32_t last_probe_m;
    int64_t csr_entries;
    int64_t reprobed_voxels;      /* FAST_LEARNED / PLAIN_EXACT exact evaluations */
    int64_t fp64_voxels;          /* voxels re-evaluated in FP64 */
    double ms_total;              /* device time, triangles resident -> field resident */
    double ms_splat;
    double ms_parity;
    double ms_fixup;
    double ms_eikonal;
    double ms_verify;
    double ms_apply;
    double ms_host_total;         /* host wall time of the call (the reference's out_time_ms) */
    double ms_sweeps;             /* device time of the directional sweep kernels alone */
    double ms_transposes;         /* device time of the x<->y plane transposes around x-sweeps */
    double ms_winding;            /* device time of the exact winding kernels alone */
    int64_t winding_pairs;        /* (voxel, triangle) pairs those kernels evaluated */
    int64_t candidates;           /* FAST_LEARNED: candidate voxels ranked */
    int64_t mismatches_found;     /* FAST_LEARNED: scanline signs the exact kernel corrected */
    double ms_proposer;           /* FAST_LEARNED: candidates + features + network + selection */
    int64_t winding_calls;        /* bracketed exact-winding launches behind ms_winding */
    int64_t sweep_calls;          /* sweep launches behind ms_sweeps (each timed on its own) */
    int64_t ranked_probes;        /* FAST_LEARNED: exact probes chosen by the scorer */
    int64_t ranked_mismatches;    /* FAST_LEARNED: scanline signs those probes corrected */
    int64_t risky_candidates;     /* FAST_LEARNED: candidates the segment certificate flags */
    int64_t fallback_probes;      /* FAST_LEARNED: flagged candidates the ranking missed */
    int64_t fallback_mismatches;  /* FAST_LEARNED: scanline signs the fallback corrected */
    int64_t far_decided;          /* exact probes decided by the far-field certificate */
    int64_t far_undecided;        /* exact probes it handed to the brute-force kernel */
    int64_t n_clusters;           /* triangle clusters of the certificate */
    double ms_far_field;          /* device time of the certificate (incl. the point sort) */
    int64_t n_parent_clusters;    /* their 2x2x2 parents (the certificate's first level) */
    int64_t far_second_pass;      /* probes the parent level left open, re-examined per cluster */
} rhz_sdf_stats;

void rhz_sdf_params_default(rhz_sdf_params* p);
const char* rhz_last_error(void);
/* The reference's fast-path log line (cuda_rasterizer.cu:1145-1157) from a
 * call's counters, into buf (cap bytes, NUL-terminated, truncated to fit):
 * "Fast SDF path OK: <total> ms total (splat+parity .. | eikonal <r> rounds ..
 * | verify+sign+copy ..) | <dims> voxels, <n> tris, band <b> vox, fixup .. |
 * verify .. probes, .. sites, .. voxels patched in .. rounds | far residual ..,
 * extra rounds .., safe-clamped ..", total = ms_host_total. */
int rhz_sdf_log_line(const rhz_sdf_stats* st, int32_t band_voxels, char* buf, int64_t cap);
/* Number of visible devices (0 when no GPU). */
int rhz_device_count(void);
/* Free and total memory of the current device.  Replaces
 * CudaRasterizer::query_gpu_memory (cuda_rasterizer.hpp:18). */
int rhz_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);
/* Launch-shape tuning of the kernels (no counterpart in the reference; used by
 * the autotuner), process-wide; results are identical for every setting:
 *   "sweep_shape_x" / "_y" / "_z"  tile each lane of an x-, y- or z-sweep walks:
 *                                  1 = one line (rows / fmajor apply), or
 *                                  10 NF + NC = NF fast-axis segments x NC cross
 *                                  lines: 21, 22, 12, 14 (defaults 14, 21, 14);
 *   "sweep_step"                   2, 4, 8: steps whose loads a sweep wave issues
 *                                  together (default 2);
 *   "sweep_rows", "sweep_fmajor"   waves per block (4, 8, 16) and 1: a block's
 *                                  waves take adjacent segments of one row, for
 *                                  the one-line walker;
 *   "fsm_layout"                   1 (default): every sweep of the rounds on a
 *                                  y-fastest copy, one plane transpose before the
 *                                  first round and one after the last; 0: the x-
 *                                  sweeps on a y-fastest copy, the others x
 *                                  fastest, two transposes every round;
 *   "sweep_run_block"              the y-sweep walking along the y-fastest rows:
 *                                  42 (default) = the prefetching kernel, 4 waves
 *                                  of one plane each per block (planes staged
 *                                  once per block, next chunk loaded while one is
 *                                  computed), adjacent plane groups in
 *                                  consecutive blocks; 41 = the same kernel, blocks
 *                                  row-group-major; 0 = the one-wave-per-plane
 *                                  kernel;
 *   "sweep_run_chunk"              16, 32 (default): cells per register chunk of
 *                                  that one-wave-per-plane kernel;
 *   "fsm_batch_rounds_max_voxels"  lattices up to this size enqueue all FSM
 *                                  rounds without host round trips;
 *   "fsm_graph"                    1: those rounds replay as one captured graph;
 *   "verify_overlap"               1 (default): on lattices of >= 16M voxels the
 *                                  FAST_LEARNED sign stage (features, ranking,
 *                                  exact probes) runs on a second stream and
 *                                  host thread while the eikonal rounds run;
 *                                  ms_proposer / ms_verify then overlap
 *                                  ms_eikonal, and ms_total is their union. */
int rhz_tuning_set(const char* key, int64_t value);
int rhz_tuning_get(const char* key, int64_t* value);
/* Per lattice-size class: "sweep_shape_x@C", "sweep_shape_y@C",
 * "sweep_shape_z@C" and "sweep_step@C" (C = 0..3) set the tile shape and the
 * steps per load group of the single builds' sweeps over lattices of class
 * C; "batch_shape_x@C", "batch_shape_y@C", "batch_shape_z@C" and
 * "batch_step@C" the same for the batched bake's sweep launches (one launch
 * per class present).  0 (the default) uses the class-free key.  Every
 * setting gives identical results; only the time differs.  No
 * counterpart in the reference (whose launch shapes are fixed): it serves
 * the autotuner's cached per-shape table. */
/* The size class of a lattice: 0 up to 2^21 voxels, 1 up to 2^24, 2 up to
 * 2^27, 3 above. */
int rhz_sweep_shape_class(const int32_t dims[3]);

/* Exhaustive self-check of the directional sweeps' square root on the current
 * device: a hardware reciprocal square root and one Newton step, compared
 * with IEEE sqrt for every binary32 argument in [2^-96, inf), the range the
 * sweeps use (each of two references counts).  *mismatches = 0 means the
 * sweeps' Godunov update rounds exactly as the reference's sqrtf.  No
 * counterpart in the reference. */
int rhz_selftest_sweep_sqrt(uint64_t* mismatches);
/* Frees the cached device workspaces of the current device: the one the
 * single-lattice calls share and the batched bake lanes' own. */
void rhz_release_workspace(void);

/* Signed distance field of triangles on the lattice [min_c, max_c] (inclusive,
 * x fastest, lattice point (min+i)*h), host buffers.
 * Replaces CudaRasterizer::compute_mesh_sdf
 * (native/src/engine/sdf/cuda_rasterizer.hpp:21-28, cuda_rasterizer.cu:1163-1265).
 * out_host holds (max-min+1) product floats; it is written only on RHZ_OK. */
int rhz_sdf_compute(const rhz_triangle* tris, int64_t n_tris, float h,
                    const int32_t min_c[3], const int32_t max_c[3],
                    const rhz_sdf_params* params, float* out_host,
                    rhz_sdf_stats* stats);

/* Same, with triangles and output already in device memory (d_out holds the
 * field on return), launched on `stream` (hipStream_t).  Replaces the device
 * part of rhi_compute_mesh_sdf_fast (cuda_rasterizer.cu:796-1159). */
int rhz_sdf_compute_device(const rhz_triangle* d_tris, int64_t n_tris, float h,
                           const int32_t min_c[3], const int32_t max_c[3],
                           const rhz_sdf_params* params, float* d_out,
                           void* stream, rhz_sdf_stats* stats);

/* One mesh of a batched bake: device triangles, lattice, device output. */
typedef struct rhz_bake_item {
    const rhz_triangle* d_tris;
    int64_t n_tris;
    float h;
    int32_t min_c[3];
    int32_t max_c[3];
    float* d_out;                 /* (max-min+1) product floats */
} rhz_bake_item;

/* Batched bake (no counterpart in the reference, whose callers build one
 * mesh per call): every item's field with the same parameters, largest
 * lattices first, on n_lanes concurrent device streams (one host thread and
 * one cached workspace per lane, 1..16).  stats[i] / ms_item[i] (host wall
 * time of item i, start to completion) are filled when non-NULL.  Returns 0
 * when every item succeeded; a declined item reports RHZ_DECLINED in
 * stats[i].status.  Results are identical to rhz_sdf_compute_device per item.
 * Ordering: the lanes start after the work already queued on `stream`
 * (hipStream_t, NULL = the legacy default stream), and the call returns when
 * every item is complete. */
int rhz_sdf_bake(const rhz_bake_item* items, int32_t n_items, const rhz_sdf_params* params, int32_t n_lanes,
                 rhz_sdf_stats* stats, double* ms_item, void* stream);

/* Indexed mesh -> triangles, dropping |(v1-v0)x(v2-v0)|^2 <= 1e-12 and
 * out-of-range indices, in index order.  Replaces gpu_build_cuda_triangles_kernel
 * + dev_try_build_cuda_triangle (cuda_rasterizer.cu:1321-1337,
 * core/math_utils.cuh:484-514).  Host buffers; *n_out receives the count. */
int rhz_build_triangles(const float* verts, int64_t n_verts, const uint32_t* idx,
                        int64_t n_tris, rhz_triangle* out, int64_t* n_out);

/* Vertex bounding box.  Replaces compute_bounding_box_gpu
 * (engine/mesh/cuda_mesh_ops.cu:197-255). */
int rhz_mesh_bbox(const float* verts, int64_t n_verts, float bmin[3], float bmax[3]);

/* build_from_mesh grid framing: bbox padded by 10 h, min = floor, max = ceil
 * (cuda_rasterizer.cu:1358-1372). */
int rhz_frame_grid(const float* verts, int64_t n_verts, float h,
                   int32_t min_c[3], int32_t max_c[3]);

/* Generalised winding number of an indexed mesh at query points.  Replaces
 * CudaRasterizer::compute_winding_numbers_gpu (cuda_rasterizer.hpp:280-285,
 * engine/mesh/cuda_mesh_ops.cu:774-891). Host buffers. */
int rhz_winding_numbers(const float* verts, int64_t n_verts, const uint32_t* idx,
                        int64_t n_tris, const float* pts, int64_t n_pts, float* out_w);

/* Winding unification and outward orientation of an indexed mesh, in place
 * on the host index buffer.  An undirected edge shared by exactly two
 * triangles links them; every connected component is unified to the winding
 * of its lowest-numbered triangle (applied when any link was traversed the
 * same way by both triangles: *had_inconsistent = 1); *is_orientable = 0 when
 * a component's links cannot all be satisfied (a Moebius band); then, if the
 * signed volume is below -1e-5, every triangle is reversed
 * (*was_inverted = 1).  *signed_volume receives the final volume (binary64,
 * fixed summation order).  Deterministic.  Out-of-range indices return
 * RHZ_ERR_ARG.  Replaces CudaRasterizer::orient_mesh_gpu
 * (cuda_rasterizer.hpp:294-301, engine/mesh/cuda_mesh_topology.cu:303-454,
 * 707-842), the pre-pass of MeshWindingPass::orient_mesh
 * (engine/io/obj_parser.cpp:8-34).  Host buffers; any output may be NULL. */
int rhz_orient_mesh(const float* verts, int64_t n_verts, uint32_t* idx, int64_t n_tris, int32_t* is_orientable,
                    int32_t* had_inconsistent, int32_t* was_inverted, double* signed_volume);
/* Signed volume sum(v0 . (v1 x v2)) / 6 (binary64, fixed order).  Replaces
 * CudaRasterizer::compute_signed_volume_gpu (cuda_rasterizer.hpp:303-306,
 * engine/mesh/cuda_mesh_topology.cu:653-705).  Host buffers. */
int rhz_signed_volume(const float* verts, int64_t n_verts, const uint32_t* idx, int64_t n_tris, double* volume);

/* Unsigned distance from each query point to an indexed mesh: the square root
 * of the smallest squared point-triangle distance over every triangle
 * (degenerate ones included).  No triangles: every distance is 0.  Replaces
 * CudaRasterizer::query_distance_batch_gpu (cuda_rasterizer.hpp:255-261,
 * engine/mesh/cuda_mesh_ops.cu:668-771). Host buffers. */
int rhz_query_distances(const float* verts, int64_t n_verts, const uint32_t* idx,
                        int64_t n_tris, const float* pts, int64_t n_pts, float* out_d);

/* Marching cubes of an x-fastest signed field at `iso`: border voxels count as
 * max(value, iso + h) (the reference's seal), a corner is inside when
 * value - iso <= 0, one vertex per cut lattice edge at origin + (i + t) h,
 * triangles from the standard table.  Writes *n_verts / *n_tris always, and
 * the vertices (3 floats each) and triangles (3 vertex ids each) when both
 * buffers are given and hold them; call with NULL buffers to size them.  The
 * order is deterministic: vertices voxel by voxel (x-, y-, z-edge), triangles
 * cube by cube.  A lattice with an axis below 2 cells has no surface.
 * Replaces CudaRasterizer::extract_iso_surface (cuda_rasterizer.hpp:73-86,
 * engine/mesh/cuda_marching_cubes.cu:829-962). Host buffers. */
int rhz_iso_surface(const float* grid, const int32_t dims[3], float voxel_size, const float origin[3],
                    float iso, float* verts, int64_t verts_cap, uint32_t* tris, int64_t tris_cap,
                    int64_t* n_verts, int64_t* n_tris);
/* The same on device buffers, on the caller's stream (hipStream_t). */
int rhz_iso_surface_device(const float* d_grid, const int32_t dims[3], float voxel_size,
                           const float origin[3], float iso, float* d_verts, int64_t verts_cap,
                           uint32_t* d_tris, int64_t tris_cap, int64_t* n_verts, int64_t* n_tris,
                           void* stream);

/* Number of values <= 0 (NaN is not inside).  The volume of the reference's
 * CudaRasterizer::compute_sdf_interior_volume_gpu (cuda_rasterizer.hpp:661-665,
 * ui/viewport/cuda_viewport_gl.cu:84-110, 702-740) is count * h^3.  Host /
 * device buffer. */
int rhz_interior_count(const float* grid, int64_t n, int64_t* count);
int rhz_interior_count_device(const float* d_grid, int64_t n, int64_t* count, void* stream);

/* Red-black Gauss-Seidel eikonal on a signed grid, `iters` iterations (the
 * reference uses 8).  Replaces CudaRasterizer::fast_sweeping_eikonal_gpu
 * (cuda_rasterizer.hpp:288-291, cuda_rasterizer.cu:1536-1629).  Host grid. */
int rhz_eikonal_redblack(float* grid, const int32_t dims[3], float h, int32_t iters);

/* ---- sign-risk proposer (no counterpart in the reference) ----------------
 * MLP 12 -> 64 -> 64 -> 64 -> 1 (ReLU), BF16 matrix-core arithmetic with FP32
 * accumulation.  Parameter blob (FP32): W1[12][64] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] w4[64] b4 (9217 trainable values) followed by the feature
 * mean[12] and inverse std[12]. */
int rhz_mlp_param_count(void);   /* 9217 */
int rhz_mlp_blob_count(void);    /* 9241 */
/* Copies the built-in trained blob (rhz_mlp_blob_count floats) to out;
 * returns 0, or RHZ_DECLINED if the library was built without one. */
int rhz_mlp_builtin(float* out);
/* logits[i] = MLP(features[i]) for n rows of 12 features (device buffers). */
int rhz_mlp_forward(const float* d_blob, const float* d_feats, int64_t n, float* d_logits, void* stream);
/* Adds the gradient of sum_i c_i * BCE(logit_i, y_i) (c_i = pos_weight for
 * y_i = 1) w.r.t. the 9217 parameters to d_grads, and the loss to *d_loss. */
int rhz_mlp_grad(const float* d_blob, const float* d_feats, const float* d_labels, int64_t n,
                 float pos_weight, float* d_grads, float* d_loss, void* stream);
/* Adam step on the first n_params values of d_blob (grad scaled by grad_scale). */
int rhz_mlp_adam(float* d_blob, const float* d_grads, float* d_m, float* d_v, int32_t n_params, float lr,
                 float beta1, float beta2, float eps, int32_t step, float grad_scale, float weight_decay,
                 void* stream);
/* Candidate voxels, their 12 features and certificate flags as the
 * FAST_LEARNED method sees them (after splat, signed scanline pass and
 * fixup): up to cap voxel ids, cap x 12 features and cap flags (1 = the
 * voxel-to-ray segment may cross a triangle) into device buffers (any may be
 * NULL); d_scan (optional, V bytes) receives the scanline signs (1 inside);
 * *n_out = total count (may exceed cap).  Features, in order: 0 nearest
 * distance / h, 1 scan sign (z ray), 2-4 nearest normal, 5 inside by a +x
 * ray, 6 planar margin, 7 inside by a +y ray (5 and 7: signed crossing
 * counts, 0.5 for a line whose crossings do not cancel), 8 log1p(triangles
 * within h), 9 normal spread within 1.5 h, 10 certificate flag, 11 distance
 * to the nearest crossing of the voxel's z ray / h (capped at 2). */
int rhz_learned_candidates(const rhz_triangle* d_tris, int64_t n_tris, float h, const int32_t min_c[3],
                           const int32_t max_c[3], const rhz_sdf_params* params, int64_t* d_voxels,
                           float* d_feats, uint8_t* d_risk, uint8_t* d_scan, int64_t cap, int64_t* n_out,
                           void* stream);

/* ---- stage-level entry points (device pointers, one stream) -------------
 * Each runs one stage of the fast path on device buffers; the pipeline above
 * is composed of exactly these.  Grids are x fastest with dims = max-min+1. */

/* [init] d[i] = value (cuda_rasterizer.cu:194-203). */
int rhz_stage_fill(float* d, int64_t n, float value, void* stream);
/* [A] exact +-band splat by atomic min (cuda_rasterizer.cu:211-268). */
int rhz_stage_splat(const rhz_triangle* d_tris, int64_t n_tris, float h, float band_width,
                    const int32_t min_c[3], const int32_t dims[3], float* d, void* stream);
/* [B] scanline parity signs (0 out, 1 in, 2 unknown) with the supercell
 * candidate lists built on the device (cuda_rasterizer.cu:279-368, :824-882).
 * *unknown_flag = 1 if any column was odd or overflowed; *csr_entries gets the
 * candidate-list size.  Returns RHZ_DECLINED if the list exceeds csr_max_entries. */
int rhz_stage_parity(const rhz_triangle* d_tris, int64_t n_tris, float h,
                     float jitter_x, float jitter_y, int32_t max_hits, int32_t sup,
                     int64_t csr_max_entries, const int32_t min_c[3], const int32_t dims[3],
                     uint8_t* d_sign, int32_t* unknown_flag, int64_t* csr_entries,
                     void* stream);
/* [C] exact winding classification of every sign==2 voxel; *fixed = count
 * (cuda_rasterizer.cu:374-400). */
int rhz_stage_fixup(const rhz_triangle* d_tris, int64_t n_tris, float h,
                    const int32_t min_c[3], const int32_t dims[3], int32_t precision,
                    uint8_t* d_sign, int64_t* fixed, void* stream);
/* [D] one directional sweep, pinned schedule: Gauss-Seidel along the swept
 * line, cross-line neighbours read from src; writes all of dst
 * (cuda_rasterizer.cu:405-513).  *changed |= 1 if any voxel decreased. */
int rhz_stage_sweep(const float* src, float* dst, const int32_t dims[3], float h,
                    int32_t axis, int32_t dir, int32_t* changed, void* stream);
/* [D] one classical fast-sweeping ordering (sx, sy, sz each +1 or -1) in place
 * on the dense field d: the lexicographic Gauss-Seidel sweep (z outermost, x
 * innermost, each axis in its direction), run as one launch per diagonal
 * level kx + ky + kz (identical results).  *changed |= 1 if any voxel
 * decreased.  No counterpart in the reference (its sweeps are the 6 axis
 * sweeps, cuda_rasterizer.cu:405-513). */
int rhz_stage_sweep_octant(float* d, const int32_t dims[3], float h, int32_t sx, int32_t sy, int32_t sz,
                           int32_t* changed, void* stream);
/* [D] one round x+, x-, y+, y-, z+, z- exactly as the pipeline runs it
 * (cuda_rasterizer.cu:1012-1017): the field is in `a` before and after; `b`
 * is the second buffer (same size). */
int rhz_stage_sweep_round(float* a, float* b, const int32_t dims[3], float h, int32_t* changed,
                          void* stream);
/* [D2] voxels >= 1e29 (cuda_rasterizer.cu:523-534). */
int rhz_stage_count_far(const float* d, int64_t n, int64_t* count, void* stream);
/* [D3] safe clamp (cuda_rasterizer.cu:544-556). */
int rhz_stage_clamp_far(float* d, const uint8_t* sign, int64_t n, float band_width,
                        int32_t intended_sign, void* stream);
/* [E] one stratified probe generation: exact winding at the 4096-style
 * sample voxels; mismatching voxel ids go to d_sites (capacity samples),
 * *m = number of mismatches (cuda_rasterizer.cu:566-607). */
int rhz_stage_verify(const rhz_triangle* d_tris, int64_t n_tris, float h,
                     const int32_t min_c[3], const int32_t dims[3], int32_t samples,
                     int32_t gen, int32_t precision, const uint8_t* d_sign,
                     int64_t* d_sites, int32_t* m, void* stream);
/* [E2] mark the (2r+1)^3 box around each site unknown (cuda_rasterizer.cu:615-646). */
int rhz_stage_mark_patch(uint8_t* d_sign, const int64_t* d_sites, int32_t n_sites,
                         int32_t radius, const int32_t dims[3], void* stream);
/* [F] signed = inside ? -d : d, d > 1e29 -> 1e6 (cuda_rasterizer.cu:650-661). */
int rhz_stage_apply(float* d, const uint8_t* sign, int64_t n, void* stream);
/* Exact winding numbers (sum / 4 pi) at lattice voxels (device list). */
int rhz_stage_winding_voxels(const rhz_triangle* d_tris, int64_t n_tris, float h,
                             const int32_t min_c[3], const int32_t dims[3],
                             const int64_t* d_voxels, int64_t n_voxels, int32_t precision,
                             double* d_w, void* stream);

/* rhz_stage_winding_voxels without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * Same kernels and results; the scratch (recompute counter, binary64
 * recompute list, split partials) is a fixed 24 MiB + 256 B device array per
 * GPU, held until the process ends, so n_voxels <= 2^20 (larger calls return
 * RHZ_ERR_ARG).  The work is enqueued on `stream` and the call returns
 * without waiting.  Calls that overlap on one GPU share that array and race:
 * order them (one stream). */
int rhz_stage_winding_voxels_noalloc(const rhz_triangle* d_tris, int64_t n_tris, float h,
                                     const int32_t min_c[3], const int32_t dims[3],
                                     const int64_t* d_voxels, int64_t n_voxels, int32_t precision,
                                     double* d_w, void* stream);

/* ---- accuracy report (no counterpart in the reference) --------------------
 * Error of a finished signed field against binary64 ground truth at n_samples
 * jittered stratified voxels (sample i in [i V / n, (i + 1) V / n), position
 * from `seed`): the distance to the nearest triangle in binary64 (brute force
 * over all triangles, lattice point as the field uses it) and the exact
 * winding sign (compensated FP32, binary64 where ||w| - 0.5| < 1e-3).
 * Errors are (|field| - distance) / h, in voxels; "band" samples are those
 * whose true distance is <= band_voxels h, the rest are "far". */
typedef struct rhz_accuracy {
    int64_t samples;
    int64_t band_samples;
    int64_t far_samples;
    int64_t unreached;         /* |field| >= 1e5 (the 1e6 marker): not in the error figures */
    int64_t on_surface;        /* true distance <= 1e-5 h: sign not checked (w = 1/2 there) */
    int64_t sign_mismatches;   /* field sign != exact winding sign (on-surface samples excluded) */
    double band_max_err;       /* max | |field| - d | / h */
    double band_rms_err;
    double band_mean_err;      /* signed mean: < 0 where the sweeps lowered band values */
    double far_max_err;
    double far_rms_err;
    double far_mean_err;       /* > 0: far distances overestimated (rounds capped) */
} rhz_accuracy;
int rhz_sdf_accuracy(const rhz_triangle* d_tris, int64_t n_tris, float h, const int32_t min_c[3],
                     const int32_t max_c[3], const float* d_field, int64_t n_samples, int32_t band_voxels,
                     uint64_t seed, void* stream, rhz_accuracy* out);

#ifdef __cplusplus
}
#endif
#endif
