// rhz_internal.hpp -- launchers of the gfx950 kernels, called by the C++ host
// orchestration (sdf_pipeline.cpp).  Every launcher enqueues on `s` and
// returns the launch status; none synchronises.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rhz {

struct Grid {
    int32_t mn[3];
    int32_t dims[3];
    float h;
    int64_t n() const { return (int64_t)dims[0] * dims[1] * dims[2]; }
};

// The lattice seen with axis `axis` (0 = x, 1 = y) rotated to the third place:
// (x, y, z) -> (y, z, x) for axis 0, (z, x, y) for axis 1.
inline Grid rotated_grid(const Grid& g, int axis) {
    const int a = axis == 0 ? 1 : 2, b = axis == 0 ? 2 : 0, c = axis == 0 ? 0 : 1;
    return Grid{{g.mn[a], g.mn[b], g.mn[c]}, {g.dims[a], g.dims[b], g.dims[c]}, g.h};
}

// Records msg as the calling thread's rhz_last_error() and returns code.
int set_error_message(int code, const char* msg);

// ---- k_grid.hip -----------------------------------------------------------
hipError_t launch_fill(float* d, int64_t n, float v, hipStream_t s);
// Zeroes n ints with a kernel (safe inside captured graphs, unlike memset nodes).
hipError_t launch_zero_i32(int* p, int n, hipStream_t s);
hipError_t launch_apply(float* d, const uint8_t* sign, int64_t n, hipStream_t s);
hipError_t launch_clamp_far(float* d, const uint8_t* sign, int64_t n, float bw, int intended,
                            hipStream_t s);
// counter[0] += number of d[i] >= 1e29
hipError_t launch_count_far(const float* d, int64_t n, unsigned long long* counter, hipStream_t s);
// The rows of an x-fastest field (len = dims[0] floats, rpp = dims[1] rows
// per plane, planes = dims[2]) copied between strides (row pitch, plane).
hipError_t launch_copy_rows(const float* src, int64_t sp, int64_t splane, float* dst, int64_t dp, int64_t dplane,
                            int len, int rpp, int planes, hipStream_t s);
// The same rows copied into the dense field d with [F] applied on the way
// (sign in the dense layout): the copy and launch_apply in one pass.
hipError_t launch_copy_apply_rows(const float* src, int64_t sp, int64_t splane, float* dst, const uint8_t* sign,
                                  int len, int rpp, int planes, hipStream_t s);
// Sentinel count over such rows (the padding between them is skipped).
hipError_t launch_count_far_rows(const float* d, int64_t p, int64_t plane, int len, int rpp, int planes,
                                 unsigned long long* counter, hipStream_t s);
// Appends i for every sign[i] == 2 to list (capacity cap); counter += count.
hipError_t launch_compact_unknown(const uint8_t* sign, int64_t n, int64_t* list, int64_t cap,
                                  unsigned long long* counter, hipStream_t s);
hipError_t launch_mark_patch(uint8_t* sign, const int64_t* sites, int n_sites, int radius,
                             const int32_t dims[3], hipStream_t s);
hipError_t launch_verify_indices(int64_t total, int samples, int gen, int64_t* out, hipStream_t s);
hipError_t launch_gather_u8(const uint8_t* src, const int64_t* idx, int64_t n, uint8_t* dst,
                            hipStream_t s);
hipError_t launch_gather_f32(const float* src, const int64_t* idx, int64_t n, float* dst, hipStream_t s);
hipError_t launch_scatter_decision(const double* w, const int64_t* idx, int64_t n, uint8_t* sign,
                                   hipStream_t s);
// sign[base + i] = |w[i]| >= 0.5
hipError_t launch_range_decision(const double* w, int64_t base, int64_t n, uint8_t* sign, hipStream_t s);
// Mismatch of sample decisions: for i < n with sign[idx[i]] != 2 and
// (sign == 1) != (|w| >= 0.5), append idx[i] to sites; counter[0] += 1.
hipError_t launch_compare_samples(const double* w, const int64_t* idx, int64_t n,
                                  const uint8_t* sign, int64_t* sites, int64_t cap,
                                  unsigned long long* counter, hipStream_t s);
// Exclusive prefix sum of n uint32 counts into n+1 int64 starts; needs
// scan_scratch_bytes(n) bytes of scratch.
size_t scan_scratch_bytes(int64_t n);
hipError_t launch_scan_u32(const uint32_t* in, int64_t n, int64_t* out, void* scratch,
                           hipStream_t s);

// ---- k_splat.hip ----------------------------------------------------------
hipError_t launch_splat(const float* tris, int64_t nt, const Grid& g, float bw, float* d,
                        hipStream_t s);

// ---- k_parity.hip ---------------------------------------------------------
struct CsrBuffers {
    float4* xyab;       // [nt]
    int4* cover;        // [nt] sx0, sx1, sy0, sy1
    uint32_t* counts;   // [nsup]
    int64_t* start;     // [nsup + 1]
    uint32_t* cursor;   // [nsup]
    int32_t* ids;       // [entries]
};
int64_t num_supercells(const Grid& g, int sup);
hipError_t launch_csr_count(const float* tris, int64_t nt, const Grid& g, int sup,
                            const CsrBuffers& b, unsigned long long* total, hipStream_t s);
hipError_t launch_csr_fill(int64_t nt, const Grid& g, int sup, const CsrBuffers& b, hipStream_t s);
// signed_count = false: reference parity; true: signed crossing count (w != 0).
hipError_t launch_parity(const float* tris, const Grid& g, int sup, const CsrBuffers& b,
                         float jx, float jy, int max_hits, uint8_t* sign, int* flag,
                         hipStream_t s, bool signed_count = false);

// ---- k_fsm.hip ------------------------------------------------------------
// Strides of a field layout in floats: between rows of the fastest axis and
// between z planes.  x-fastest: (x, y, z) at z plane + y pitch + x;
// y-fastest: (x, y, z) at z plane + x pitch + y.
struct FieldLayout {
    int64_t pitch;
    int64_t plane;
};
inline FieldLayout dense_xy(const Grid& g) { return {g.dims[0], (int64_t)g.dims[0] * g.dims[1]}; }
inline FieldLayout dense_yx(const Grid& g) { return {g.dims[1], (int64_t)g.dims[0] * g.dims[1]}; }
// Sweep of an x-fastest field (dense, or layout L).  Axis 0 (generic
// kernels) takes dense fields only.
hipError_t launch_sweep(const float* src, float* dst, const Grid& g, int axis, int dir,
                        int* changed, hipStream_t s);
hipError_t launch_sweep(const float* src, float* dst, const Grid& g, int axis, int dir,
                        int* changed, hipStream_t s, const FieldLayout& L);
// True when the lane-walker sweeps apply (every axis has >= 2 cells, a plane
// has < 2^27 cells, 2^-30 <= h <= 2^30); otherwise launch_sweep uses generic kernels.
bool lane_sweeps_ok(const Grid& g);
struct SweepTuning {
    int step;  // walk steps whose loads are issued together (2, 4, 8)
    int rows;  // waves per block (4, 8, 16)
    int fmajor;  // 1: a block's waves are adjacent segments of one row; 0: adjacent rows
    // Tile shape per sweep orientation (x: on the y-fastest copy; y; z): 1 =
    // one-line walker (rows / fmajor apply); 10 NF + NC = tile of NF fast-axis
    // segments x NC cross lines per lane, one of 21, 22, 12, 14.
    int shape[3];
    int64_t batch_rounds_max_voxels;  // lattices up to this size enqueue all FSM rounds at once
    int fsm_graph;                    // 1: those rounds are captured and replayed as one hipGraph
    // Per lattice-size class (shape_class): tile shape per orientation and
    // steps per load group; 0 = the class-free value above.
    int cls_shape[4][3];
    int cls_step[4];
    // The same for the batched bake's sweep launches (a stage of its own: they
    // share the GPU with the sign stages); 0 = the class-free value.
    int bcls_shape[4][3];
    int bcls_step[4];
    int run_chunk;   // cells per register chunk of sweep_run_kernel (16, 32)
    int fsm_layout;  // 0: x-sweeps on a y-fastest copy, the rest x-fastest (two transposes
                     // a round); 1: every sweep on the y-fastest copy (no transposes between rounds)
    int run_block;   // y-sweep on the y-fastest copy: 0 = sweep_run_kernel, 41 = the prefetching
                     // kernel (4 waves of one plane each per block), 42 = the same with
                     // plane-group-major block order (the default)
};
// Lattice-size classes of the per-class sweep tuning: <= 2^21, <= 2^24,
// <= 2^27 voxels, larger.
constexpr int kShapeClasses = 4;
int shape_class(int64_t voxels);
// The tile shape (orientation `axis`: the walk axis) and steps per load group
// a sweep over a lattice of `voxels` cells uses.
int effective_shape(int axis, int64_t voxels);
int effective_step(int64_t voxels);
SweepTuning sweep_tuning();
bool set_sweep_tuning(int step, int rows, int fmajor);
bool set_sweep_shape(int axis, int shape);
// shape / step 0 clears the class's own value.
bool set_class_shape(int cls, int axis, int shape);
bool set_class_step(int cls, int step);
bool set_run_chunk(int cells);
bool set_run_block(int shape);
bool set_fsm_layout(int layout);
bool set_batch_class_shape(int cls, int axis, int shape);
bool set_batch_class_step(int cls, int step);
void set_fsm_batching(int64_t max_voxels, int graph);
// One classical fast-sweeping ordering (sx, sy, sz in {+1, -1}) in place on
// the dense x-fastest field g: the lexicographic Gauss-Seidel sweep, run as
// one launch per diagonal level.  *changed |= 1 if a voxel decreased.
hipError_t launch_sweep_octant(float* g, const Grid& gr, int sx, int sy, int sz, int* changed, hipStream_t s);
// Per-plane transpose between the x-fastest and the y-fastest layouts (dense,
// or the layouts of the input and the output).
hipError_t launch_transpose_xy(const float* in, float* out, const Grid& g, bool to_yx, hipStream_t s);
// y-fastest field (layout lin) -> dense x-fastest field with the [F] sign apply
hipError_t launch_transpose_apply(const float* in, const FieldLayout& lin, float* out, const uint8_t* sign,
                                  const Grid& g, hipStream_t s, unsigned long long* far = nullptr);
// Exhaustive device check of the sweeps' square root: mismatches against IEEE
// sqrt over every binary32 argument in [2^-96, inf) are added to *bad.
hipError_t launch_sqrt_selftest(unsigned long long* bad, hipStream_t s);
hipError_t launch_transpose_xy(const float* in, float* out, const Grid& g, bool to_yx, hipStream_t s,
                               const FieldLayout& lin, const FieldLayout& lout);
// Any sweep on the y-fastest layout L (rows padded to 32 floats for axis 1):
// x as launch_sweep_yx, y walking along the rows (sweep_run_kernel), z with
// lanes over y; results equal the x-fastest sweeps bit for bit.
hipError_t launch_sweep_on_yx(const float* src, float* dst, const Grid& g, int axis, int dir, int* changed,
                              hipStream_t s, const FieldLayout& L);
// x-sweep on the y-fastest layout (src/dst both y-fastest; dense, or layout L).
hipError_t launch_sweep_yx(const float* src, float* dst, const Grid& g, int dir, int* changed, hipStream_t s);
hipError_t launch_sweep_yx(const float* src, float* dst, const Grid& g, int dir, int* changed, hipStream_t s,
                           const FieldLayout& L);
// The padded layouts of the FSM's working copies (x-fastest and y-fastest).
void fsm_layouts(const Grid& g, FieldLayout* xy, FieldLayout* yx);

// ---- batched rounds (many lattices per launch) ----------------------------
// One lattice's part of a batched directional sweep: the lane sweep's
// arguments (fast axis nf cells, walk nw steps of stride ws, nc cross lines
// of stride cs) and the lattice's per-round change flags (flags[r] = round r
// lowered a voxel; round r > 0 runs only if flags[r - 1] != 0).
struct SweepJob {
    const float* src;
    float* dst;
    int64_t ws, cs;
    int nf, nw, nc, nbf;
    float h;
    int* flags;
};
// One lattice's part of a batched plane transpose (gated like the sweeps).
struct TransJob {
    const float* in;
    float* out;
    int64_t in_pitch, in_plane, out_pitch, out_plane;
    int rows, cols, tiles_r, tiles_c;
    int* flags;
};
// Sweep orientations: 0 = x-sweep on the y-fastest copy, 1 = y-sweep, 2 =
// z-sweep (x-fastest copy).  Blocks of one job, and its fast-axis block
// count (the job's nbf), for the batched tile shape of that orientation and
// lattice-size class; one launch covers jobs of one class.
int64_t sweep_batch_blocks(int orient, int cls, int nf, int nc, int* nbf);
hipError_t launch_sweep_batch(int orient, int cls, int dir, const SweepJob* jobs, const unsigned* starts, int njobs,
                              unsigned nblocks, int r, hipStream_t s);
int64_t transpose_batch_blocks(int rows, int cols, int planes, int* tiles_r, int* tiles_c);
hipError_t launch_transpose_batch(const TransJob* jobs, const unsigned* starts, int njobs, unsigned nblocks, int r,
                                  hipStream_t s);
// ---- k_mc.hip -------------------------------------------------------------
// Marching cubes in three passes: classify (in: inside byte per voxel; info: u8 per voxel; blk_v / blk_t:
// u32 per block of mc_blocks(dims)), prefix sums of the block totals into
// vstart / tstart (int64, blocks + 1 entries), then emission (voff: u32 per
// voxel scratch; verts: 3 floats per vertex; tris: 3 ids per triangle).
int64_t mc_blocks(const int32_t dims[3]);
hipError_t launch_mc_classify(const float* g, const int32_t dims[3], float h, float iso, uint8_t* in, uint8_t* info,
                              uint32_t* blk_v, uint32_t* blk_t, hipStream_t s);
hipError_t launch_mc_emit(const float* g, const int32_t dims[3], float h, const float origin[3], float iso,
                          const uint8_t* in, const uint8_t* info, const int64_t* vstart, const int64_t* tstart,
                          uint32_t* voff, float* verts, uint32_t* tris, hipStream_t s);
// count += number of values <= 0 among g[0..n).
hipError_t launch_interior_count(const float* g, int64_t n, unsigned long long* count, hipStream_t s);
hipError_t launch_redblack(float* g, const int32_t dims[3], float h, int parity, hipStream_t s);

// ---- k_winding.hip --------------------------------------------------------
// Winding numbers w = sum/(4 pi) at lattice voxels idx[0..n) (or base + i
// when idx is null).  precision 0 = plain FP32 in triangle order,
// 1 = compensated FP32 with a binary64 recompute where ||w| - 0.5| < fp64_band.
// part: split partials scratch (winding_scratch_bytes), fp64_list/fp64_cnt:
// recompute list (capacity n) and its counter.
size_t winding_scratch_bytes(int64_t n_points, int64_t n_tris);
hipError_t launch_winding(const float* tris, int64_t nt, const Grid& g, const int64_t* idx,
                          int64_t base, int64_t n, int precision, float fp64_band, double* w,
                          void* scratch, int64_t* fp64_list, unsigned long long* fp64_cnt,
                          hipStream_t s);
// Second half of launch_winding's binary64 step (needs the count on the host).
hipError_t launch_winding_fp64(const float* tris, int64_t nt, const Grid& g, const int64_t* idx,
                               int64_t base, const int64_t* fp64_list, int64_t n_fp64, double* w,
                               hipStream_t s);
// Binary64 squared distance from lattice voxels idx[0..n) to the nearest of
// the nt triangles (brute force; +inf when nt = 0).
hipError_t launch_distance_f64(const float* tris, int64_t nt, const Grid& g, const int64_t* idx, int64_t n,
                               double* d2, hipStream_t s);
// The same with the recompute count read on the device from *count (the grid
// covers cap list entries): no host round trip.
hipError_t launch_winding_fp64_count(const float* tris, int64_t nt, const Grid& g, const int64_t* idx,
                                     int64_t base, const int64_t* fp64_list, const unsigned long long* count,
                                     int64_t cap, double* w, hipStream_t s);
// Legacy exact per-voxel kernel (distance + winding sign).
hipError_t launch_legacy(const float* tris, int64_t nt, const Grid& g, float* out, void* scratch,
                         hipStream_t s);
size_t legacy_scratch_bytes(int64_t nt);
// Standalone winding numbers over an indexed mesh.
hipError_t launch_winding_points(const float* verts, const uint32_t* idx, int64_t nt,
                                 const float* pts, int64_t np, float* out, hipStream_t s);
// Unsigned distance of each point to an indexed mesh (brute force).
hipError_t launch_point_distances(const float* verts, const uint32_t* idx, int64_t nt, const float* pts, int64_t np,
                                  float* out, hipStream_t s);

// ---- k_learned.hip --------------------------------------------------------
// Candidate voxels: band distance d[i] <= thr (appended to list, counter += n).
// k_cluster: certified far-field winding decisions
// Cells of side 1 / inv_s from (ox, oy, oz); nx x ny x nz cells, indexed
// parent-major: the 2 x 2 x 2 cells of parent (X, Y, Z) are the eight
// consecutive ids 8 ((Z NY + Y) NX + X) + (x & 1) + 2 (y & 1) + 4 (z & 1),
// NX = ceil(nx / 2) and so on; cells() = 8 NX NY NZ.
struct ClusterGrid {
    float ox, oy, oz, inv_s;
    int nx, ny, nz;
    __host__ __device__ int px() const { return (nx + 1) / 2; }
    __host__ __device__ int py() const { return (ny + 1) / 2; }
    __host__ __device__ int pz() const { return (nz + 1) / 2; }
    __host__ __device__ int64_t parents() const { return (int64_t)px() * py() * pz(); }
    __host__ __device__ int64_t cells() const { return 8 * parents(); }
};
hipError_t launch_tri_box_area(const float* tris, int64_t nt, int* box, double* area_part, int nblocks,
                               hipStream_t s);
// Clusters of the non-empty cells in id order (info: 5 float4 each; range:
// first triangle id slot and count) and of the non-empty parents (pinfo,
// prange: first child cluster and count); slot: cells() + 1 int64, pslot:
// parents() + 1 int64.  The cluster counts are slot[cells()] and
// pslot[parents()].
hipError_t launch_cluster_build(const float* tris, int64_t nt, const ClusterGrid& g, uint32_t* counts,
                                uint32_t* cursor, int64_t* start, void* scan_scratch, int32_t* ids, float4* info,
                                int2* range, int64_t* slot, uint32_t* pflags, int64_t* pslot, float4* pinfo,
                                int2* prange, hipStream_t s);
// Spatial order of list positions (pos: the m positions to sort, or NULL for
// 0..m-1) by cluster-grid cell, into order[m].
hipError_t launch_point_sort(const Grid& g, const int64_t* idx, const int64_t* pos, int64_t m, const ClusterGrid& cg,
                             uint32_t* counts, uint32_t* cursor, int64_t* start, void* scan_scratch, int64_t* order,
                             hipStream_t s);
// Certified winding numbers of the n list positions order[0..n) (voxel ids in
// idx); undecided positions are appended to `undecided`.  use_parents = false
// evaluates every cluster at the child level (tighter bounds).  The loop over
// the npar parent clusters is split into far_parts(n, npar) ranges run as
// separate blocks; scratch holds their partial sums (far_scratch_bytes(n,
// npar) bytes).
int far_parts(int64_t n, int ncl);
size_t far_scratch_bytes(int64_t n, int ncl);
hipError_t launch_winding_far(const float* tris, const float4* info, const int2* range, const int32_t* ids,
                              const float4* pinfo, const int2* prange, int npar, bool use_parents, const Grid& g,
                              const int64_t* idx, const int64_t* order, int64_t n, float beta, void* scratch,
                              double* w, int64_t* undecided, unsigned long long* n_undecided, hipStream_t s);
hipError_t launch_gather_i64(const int64_t* src, const int64_t* pos, int64_t m, int64_t* out, hipStream_t s);
hipError_t launch_scatter_f64(const double* src, const int64_t* pos, int64_t m, double* dst, hipStream_t s);

hipError_t launch_candidates(const float* d, int64_t n, float thr, int64_t* list, int64_t cap,
                             unsigned long long* counter, hipStream_t s);
// The column candidate lists in (lowest z, triangle id) order for the
// features: zr = every triangle's z extent, *max_ext = the largest, keys =
// the sorted (zkey << 32 | id) entries (keys_tmp: the unsorted ones);
// scratch: zsort_scratch_bytes.
size_t zsort_scratch_bytes(int64_t entries, int64_t nsup, const int64_t* start);
hipError_t launch_zsort_lists(const float* tris, int64_t nt, const CsrBuffers& b, int64_t entries, int64_t nsup,
                              float2* zr, unsigned* max_ext, unsigned long long* keys_tmp, unsigned long long* keys,
                              void* scratch, size_t scratch_bytes, hipStream_t s);
// 12 features per candidate (row-major n x 12) from the z-sorted lists.
hipError_t launch_features(const float* tris, int64_t nt, const CsrBuffers& b, const float2* zr,
                           const unsigned long long* keys, const unsigned* max_ext, const uint8_t* sign,
                           const int64_t* list, int64_t n, const Grid& g, int sup, float jx, float jy, float* feats,
                           uint8_t* risk, unsigned long long* n_risky, hipStream_t s);
hipError_t launch_unprobed_risky(const int64_t* list, const uint8_t* risk, int64_t n, const unsigned* bits,
                                 int64_t* out, int64_t cap, unsigned long long* counter, hipStream_t s);
hipError_t launch_margin_score(const float* feats, int64_t n, float* score, hipStream_t s);
// Triangles with axis `axis` (0 = x, 1 = y) rotated to the third coordinate
// (as rotated_grid), for column lists along that axis.
hipError_t launch_rotate_tris(const float* tris, int64_t nt, int axis, float* out, hipStream_t s);
// Row lists of the rotated lattice r = rotated_grid(g, axis): row (iu, iv)
// holds the rotated triangles whose (u, v) box contains its jittered ray
// point.  Count (cover: int4 per triangle; counts: u32 per row; big: int32
// per triangle, with its count n_big, for the triangles of many rows; *total
// += entries), then, after the exclusive scan of counts into start, fill.
hipError_t launch_axis_rows_count(const float* rtris, int64_t nt, const Grid& r, float jx, float jy, int4* cover,
                                  uint32_t* counts, int32_t* big, unsigned long long* n_big,
                                  unsigned long long* total, hipStream_t s);
hipError_t launch_axis_rows_fill(int64_t nt, const Grid& r, const int4* cover, const int32_t* big,
                                 const unsigned long long* n_big, const int64_t* start, uint32_t* cursor, int32_t* ids,
                                 hipStream_t s);
// Features `slot` of the n candidates: inside by the signed crossings of the
// ray along `axis` (1), outside (0), open line (0.5).
hipError_t launch_axis_parity(const float* rtris, const int64_t* start, const int32_t* ids, const int64_t* list,
                              int64_t n, const Grid& g, int axis, float jx, float jy, float* feats, int slot,
                              hipStream_t s);
hipError_t launch_fill_feature(float* feats, int64_t n, int slot, float value, hipStream_t s);
// A uniform position in [0, 0.5) hashed from each candidate's voxel index, plus
// the certificate flag (feature 10) when `feats` is not null.
hipError_t launch_hash_score(const float* feats, const int64_t* cand, int64_t n, float* score, hipStream_t s);
int score_bins();
hipError_t launch_score_hist(const float* score, int64_t n, float lo, float hi, unsigned* hist, hipStream_t s);
hipError_t launch_select(const float* score, const int64_t* list, int64_t n, float lo, float hi, int min_bin,
                         int64_t* out, int64_t cap, unsigned long long* counter, hipStream_t s);
hipError_t launch_mark_probed(const int64_t* list, int64_t n, unsigned* bits, hipStream_t s);
// sign[v] = exact decision; voxels whose previous sign differed go to mism.
hipError_t launch_commit(const double* w, const int64_t* list, int64_t n, uint8_t* sign, int64_t* mism,
                         int64_t cap, unsigned long long* counter, hipStream_t s);
// Unprobed 26-neighbours of the mismatches (marked probed as they are claimed).
hipError_t launch_grow(const int64_t* mism, int64_t m, const int32_t dims[3], unsigned* bits, int64_t* out,
                       int64_t cap, unsigned long long* counter, hipStream_t s);

// ---- k_mlp.hip ------------------------------------------------------------
int mlp_param_count();  // 9217 trainable
int mlp_blob_count();   // + feature mean / inverse std
hipError_t launch_mlp_forward(const float* P, const float* X, int64_t n, float* out, hipStream_t s);
hipError_t launch_mlp_grad(const float* P, const float* X, const float* Y, int64_t n, float pos_weight, float* G,
                           float* loss, hipStream_t s);
hipError_t launch_adam(float* P, const float* G, float* M, float* V, int n, float lr, float b1, float b2, float eps,
                       int step, float scale, float wd, hipStream_t s);

// ---- k_topology.hip -------------------------------------------------------
// Open-addressing table of undirected edges (size a power of two, at least
// twice the edge count: edge_table_size(nt)); per key, the number of
// triangles on the edge and the min / max of (triangle << 1 | direction).
struct EdgeTable {
    unsigned long long* key;
    uint32_t* count;
    unsigned long long* lo;
    unsigned long long* hi;
    int64_t size;
};
int64_t edge_table_size(int64_t nt);
// Edge table of idx, then the links: nb[3 t + e] = the other triangle on edge
// e of t when exactly two share it (else -1), same[3 t + e] = 1 when both
// traverse it the same way (then *inconsistent = 1).  slot: 3 nt int64.
hipError_t launch_edge_links(const uint32_t* idx, int64_t nt, const EdgeTable& E, int64_t* slot, int32_t* nb,
                             uint8_t* same, int* inconsistent, hipStream_t s);
hipError_t launch_label_init(int64_t nt, int32_t* parent, uint8_t* parity, hipStream_t s);
// One hook (p0, q0 -> p1, q1) and one jump (p1, q1 -> p0, q0); *changed = 1
// if a parent moved.
hipError_t launch_label_step(int64_t nt, const int32_t* nb, const uint8_t* same, int32_t* p0, uint8_t* q0,
                             int32_t* p1, uint8_t* q1, int* changed, hipStream_t s);
hipError_t launch_orient_check(int64_t nt, const int32_t* nb, const uint8_t* same, const uint8_t* parity,
                               int* non_orientable, hipStream_t s);
// Reverses triangles with flip[t] != 0 (all when flip is null).
hipError_t launch_flip(uint32_t* idx, int64_t nt, const uint8_t* flip, hipStream_t s);
// volume_blocks(nt) binary64 partial sums of v0 . (v1 x v2).
int volume_blocks(int64_t nt);
hipError_t launch_volume_partials(const float* v, const uint32_t* idx, int64_t nt, double* part, hipStream_t s);

// ---- k_mesh.hip -----------------------------------------------------------
hipError_t launch_tri_valid(const float* verts, int64_t nv, const uint32_t* idx, int64_t nt,
                            uint32_t* valid, hipStream_t s);
hipError_t launch_tri_emit(const float* verts, const uint32_t* idx, int64_t nt,
                           const uint32_t* valid, const int64_t* pos, float* out, hipStream_t s);
hipError_t launch_bbox(const float* verts, int64_t nv, float* partial, int nblocks, float* out6,
                       hipStream_t s);
int bbox_blocks(int64_t nv);

}  // namespace rhz
