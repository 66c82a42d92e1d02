// sdf_pipeline.cpp -- C++ host orchestration of the mesh-to-SDF path and the
// C-ABI (include/rhizome_sdf.h).
//
// The policy is the reference rhi_compute_mesh_sdf_fast
// (native/src/engine/sdf/cuda_rasterizer.cu:796-1159) and compute_mesh_sdf
// (:1163-1265): guards, candidate lists, fill, splat, parity, exact fixup, FSM
// rounds with sentinel audit / audit-gated extension / safe clamp, stratified
// exact-winding verification with patch boxes, sign apply, and the 32Mi gate
// with the legacy fallback.  Differences by design:
//   * device memory comes from a per-device grow-only workspace instead of
//     eight cudaMalloc/cudaFree per call;
//   * candidate lists are built on the device;
//   * every stage is enqueued on one stream; the host reads back only the
//     small counters the policy branches on (as the reference does);
//   * the FSM uses a fixed sweep schedule (see k_fsm.hip) and two buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <thread>
#include <climits>
#include <string>
#include <vector>

#include "rhizome_sdf.h"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

constexpr float kFar = 1e30f;  // unreached marker (cuda_rasterizer.cu:77)

thread_local std::string g_err;

int set_err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace

int set_error_message(int code, const char* msg) { return set_err(code, "%s", msg); }

namespace {

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return set_err(RHZ_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e));       \
    } while (0)

// ---------------------------------------------------------------------------
// per-device workspace
// ---------------------------------------------------------------------------
enum Slot {
    S_DIST2 = 0,   // FSM working buffer
    S_SIGN,        // sign volume
    S_XYAB,
    S_COVER,
    S_COUNTS,
    S_START,
    S_CURSOR,
    S_IDS,
    S_SCAN,
    S_LIST,        // voxel index list
    S_W,           // winding numbers (double)
    S_WSCR,        // winding split partials
    S_FP64,        // binary64 recompute list
    S_SITES,
    S_SAMPLES,
    S_TRIS,        // triangles staged by the host API
    S_OUT,         // field staged by the host API
    S_MISC,
    S_MISC2,
    S_BITS,        // probed-voxel bitmap (learned method)
    S_CAND,        // candidate voxel list
    S_FEAT,        // candidate features
    S_SCORE,       // candidate scores
    S_SEL,         // selected / grown probe list
    S_MISM,        // mismatch list
    S_HIST,        // score histogram
    S_MLP,         // proposer parameters (device copy)
    S_RISK,        // per-candidate segment-certificate flag
    S_CL_BOX,      // triangle box (ordered ints) + per-block area partials
    S_CL_COUNTS,   // triangles per cluster cell
    S_CL_CURSOR,
    S_CL_START,
    S_CL_SCAN,
    S_CL_IDS,      // triangle ids grouped by cell
    S_CL_INFO,     // per cluster: centre, radius, area vector, area
    S_CL_RANGE,    // per cluster: first id, count
    S_CL_SLOT,     // cluster slot of every cell (prefix sums of non-empty flags)
    S_CL_PFLAG,    // non-empty parents
    S_CL_PSLOT,    // parent slots
    S_CL_PINFO,    // per parent cluster: centre, radius, area vector, area, moment
    S_CL_PRANGE,   // per parent cluster: first child cluster, count
    S_UNDEC,       // positions the far-field certificate left undecided
    S_UNDEC2,      // ... and still undecided at the child level
    S_UNDVOX,      // their voxel ids
    S_W2,          // their brute-force winding numbers
    S_ORDER,       // probe positions in spatial order
    S_FARPART,     // far-field partial sums per cluster range
    S_PROBES,      // ranked probes followed by the certificate fallback
    S_ZRANGE,      // z extent of every triangle (candidate features)
    S_MC_INFO,     // marching cubes: cut mask + triangle count per voxel
    S_MC_IN,       // marching cubes: inside byte per voxel
    S_MC_BLK,      // per-block vertex and triangle totals
    S_MC_START,    // their prefix sums
    S_MC_SCAN,
    S_MC_VOFF,     // first vertex id per voxel
    S_MC_VERTS,    // vertices / triangles staged by the host API
    S_MC_TRIS,
    S_DIST3,       // second FSM working buffer (rows padded to 128 B)
    S_ACC,         // accuracy report: field values at the samples
    S_TP_KEY,      // orientation: edge table keys, counts, min / max tags
    S_TP_CNT,
    S_TP_LO,
    S_TP_HI,
    S_TP_SLOT,     // table slot of every triangle edge
    S_TP_NB,       // linked triangle of every edge, and whether it agrees
    S_TP_SAME,
    S_TP_P0,       // component labels (double-buffered) and parities
    S_TP_Q0,
    S_TP_P1,
    S_TP_Q1,
    S_TP_VOL,      // signed-volume partial sums
    S_ZKEY,        // candidate lists sorted by (lowest z, id), unsorted keys, sort scratch, largest z extent
    S_ZKEY_TMP,
    S_ZSORT,
    S_ZEXT,
    S_AX_TRIS,     // triangles rotated for x / y rays, and their row lists
    S_AX_COVER,
    S_AX_COUNTS,
    S_AX_START,
    S_AX_CURSOR,
    S_AX_IDS,
    S_AX_SCAN,
    S_AX_BIG,      // triangles covering many rows (handled a wave each)
    S_COUNT
};

constexpr int kNumCounters = 32;

struct Workspace {
    void* p[S_COUNT] = {};
    size_t bytes[S_COUNT] = {};
    unsigned long long* d_cnt = nullptr;  // device counters (ull[16] + int[32])
    unsigned long long* h_cnt = nullptr;  // pinned mirror
    bool init = false;
    // the sign stage's own workspace and stream while it overlaps the
    // eikonal rounds (created on first use)
    Workspace* side = nullptr;
    hipStream_t side_stream = nullptr;
    // buffers a slot outgrew: hipFree waits for the whole device, so they are
    // freed by ws_reap between calls, never while a call's work is queued
    std::vector<void*> retired;
};

// Frees the buffers that w and its side workspace retired.
void ws_reap(Workspace* w) {
    if (!w) return;
    for (Workspace* x : {w, w->side}) {
        if (!x) continue;
        for (void* p : x->retired) (void)hipFree(p);
        x->retired.clear();
    }
}

std::mutex g_ws_mu;
Workspace g_ws[64];

int current_device() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return -1;
    return d;
}

Workspace* workspace() {
    int d = current_device();
    if (d < 0 || d >= 64) return nullptr;
    Workspace& w = g_ws[d];
    if (!w.init) {
        if (hipMalloc(&w.d_cnt, 512) != hipSuccess) return nullptr;
        if (hipHostMalloc((void**)&w.h_cnt, 512, hipHostMallocDefault) != hipSuccess) return nullptr;
        w.init = true;
    }
    ws_reap(&w);  // the previous call's retired buffers (callers hold g_ws_mu)
    return &w;
}

// Batched bake lanes: each lane owns a workspace and a stream (default
// priority, non-blocking) and is driven by one host thread.
constexpr int kMaxLanes = 16;
struct Lane {
    Workspace ws;
    hipStream_t stream = nullptr;
    bool ready = false;
};
std::mutex g_lane_mu;
Lane g_lanes[64][kMaxLanes];

// Device memory of the batched rounds, cached per device (grow-only).
struct BakeWork {
    Workspace ws;            // slots below, plus counters
    hipStream_t stream = nullptr;
    bool ready = false;
};
BakeWork g_bake[64];

Lane* bake_lane(int lane) {
    int d = current_device();
    if (d < 0 || d >= 64 || lane < 0 || lane >= kMaxLanes) return nullptr;
    Lane& L = g_lanes[d][lane];
    if (!L.ready) {
        if (hipMalloc(&L.ws.d_cnt, 512) != hipSuccess) return nullptr;
        if (hipHostMalloc((void**)&L.ws.h_cnt, 512, hipHostMallocDefault) != hipSuccess) return nullptr;
        if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess) return nullptr;
        L.ws.init = true;
        L.ready = true;
    }
    return &L;
}

// The side workspace of `w`: own counters, mirror, buffers and a stream
// (default priority, non-blocking).
Workspace* side_of(Workspace* w) {
    if (w->side) return w->side;
    Workspace* sw = new Workspace;
    if (hipMalloc(&sw->d_cnt, 512) != hipSuccess ||
        hipHostMalloc((void**)&sw->h_cnt, 512, hipHostMallocDefault) != hipSuccess ||
        hipStreamCreateWithFlags(&w->side_stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        if (sw->d_cnt) (void)hipFree(sw->d_cnt);
        if (sw->h_cnt) (void)hipHostFree(sw->h_cnt);
        delete sw;
        w->side_stream = nullptr;
        return nullptr;
    }
    sw->init = true;
    w->side = sw;
    return sw;
}

// Returns a device buffer of at least `bytes` for `slot` (grow-only).  A slot
// that has to grow takes at least 1.5x its old size (a bake lane meets items
// of many sizes), and the buffer it outgrew is retired, not freed: a hipFree
// here would wait for every stream on the device, the other bake lanes and
// the side stream included.
void* ws_get(Workspace* w, int slot, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (w->bytes[slot] >= bytes) return w->p[slot];
    const size_t old = w->bytes[slot];
    if (w->p[slot]) {
        w->retired.push_back(w->p[slot]);
        w->p[slot] = nullptr;
        w->bytes[slot] = 0;
    }
    const size_t exact = (bytes + 4095) & ~size_t(4095);
    if (old > 0) bytes = std::max(bytes, old + old / 2);
    size_t rounded = (bytes + 4095) & ~size_t(4095);
    hipError_t e = hipMalloc(&w->p[slot], rounded);
    if (e != hipSuccess && !w->retired.empty()) {
        // out of memory: give the retired buffers back first (hipFree waits
        // for the device, so the work still queued on them completes)
        (void)hipGetLastError();
        for (void* p : w->retired) (void)hipFree(p);
        w->retired.clear();
        e = hipMalloc(&w->p[slot], rounded);
    }
    if (e != hipSuccess && rounded > exact) {
        // the 1.5x growth margin does not fit: take exactly what was asked
        (void)hipGetLastError();
        rounded = exact;
        e = hipMalloc(&w->p[slot], rounded);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        w->p[slot] = nullptr;
        return nullptr;
    }
    w->bytes[slot] = rounded;
    return w->p[slot];
}

// counters: ull slots 0..15, int slots at byte offset 128 (32 ints)
enum CounterId {
    C_CSR = 0, C_LIST = 1, C_FP64 = 2, C_FAR = 3, C_MISM = 4, C_SEL = 5, C_GROW = 6, C_RISKY = 7, C_NCL = 8,
    C_UNDEC = 9, C_MISM2 = 10, C_INSIDE = 11, C_UNDEC2 = 12, C_AXCSR = 13, C_AXBIG = 14
};
enum FlagId { F_UNKNOWN = 0, F_CHANGED = 1, F_ROUND0 = 2 };  // F_ROUND0 + r: round r changed a voxel
constexpr int kMaxBatchedRounds = 24;

unsigned long long* cnt_ptr(Workspace* w, int id) { return w->d_cnt + id; }
int* flag_ptr(Workspace* w, int id) { return reinterpret_cast<int*>(reinterpret_cast<char*>(w->d_cnt) + 128) + id; }

struct Timer {
    hipEvent_t ev[8] = {};
    int n = 0;
    bool ok = false;
    explicit Timer(bool enable) {
        if (!enable) return;
        ok = true;
        for (int i = 0; i < 8; ++i)
            if (hipEventCreate(&ev[i]) != hipSuccess) ok = false;
    }
    ~Timer() {
        for (int i = 0; i < 8; ++i)
            if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
    void mark(int i, hipStream_t s) {
        if (ok) (void)hipEventRecord(ev[i], s);
    }
    double ms(int a, int b) {
        float t = 0.f;
        if (!ok || hipEventElapsedTime(&t, ev[a], ev[b]) != hipSuccess) return 0.0;
        return t;
    }
};

// Accumulates device time of bracketed launch sequences; events are read
// once, after the caller's final synchronisation.
struct EventAcc {
    std::vector<hipEvent_t> evs;
    size_t used = 0;
    ~EventAcc() {
        for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    }
    hipEvent_t next() {
        if (used == evs.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            evs.push_back(e);
        }
        return evs[used++];
    }
    void begin(hipStream_t s) {
        hipEvent_t e = next();
        if (e) (void)hipEventRecord(e, s);
    }
    void end(hipStream_t s) { begin(s); }
    double total_ms() {
        double t = 0.0;
        for (size_t i = 0; i + 1 < used; i += 2) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, evs[i], evs[i + 1]) == hipSuccess) t += ms;
        }
        return t;
    }
};

struct Ctx {
    const float* tris;
    int64_t nt;
    Grid g;
    const rhz_sdf_params* p;
    float* d;  // field buffer (device), V floats
    hipStream_t s;
    rhz_sdf_stats* st;
    Workspace* w;
    EventAcc* sweeps_t = nullptr;
    EventAcc* wind_t = nullptr;
    EventAcc* trans_t = nullptr;
    EventAcc* far_t = nullptr;
    // far-field clusters of this call's triangles (built on first use)
    int ncl = -1;
    ClusterGrid cl_grid{};
    const float4* cl_info = nullptr;
    const int2* cl_range = nullptr;
    const float4* cl_pinfo = nullptr;
    const int2* cl_prange = nullptr;
    int npar = 0;
    const int32_t* cl_ids = nullptr;
};

// Reads `count` ull counters starting at `first` into h_cnt (synchronises).
int read_counters(Ctx& c) {
    HIP_TRY(hipMemcpyAsync(c.w->h_cnt, c.w->d_cnt, 512, hipMemcpyDeviceToHost, c.s));
    HIP_TRY(hipStreamSynchronize(c.s));
    return RHZ_OK;
}
unsigned long long hc(Ctx& c, int id) { return c.w->h_cnt[id]; }
int hf(Ctx& c, int id) { return reinterpret_cast<int*>(reinterpret_cast<char*>(c.w->h_cnt) + 128)[id]; }

int zero_counters(Ctx& c) {
    HIP_TRY(hipMemsetAsync(c.w->d_cnt, 0, 512, c.s));
    return RHZ_OK;
}

// Exact winding decisions for `n` voxels (explicit list or base + i), written
// as signs 0/1 into `sign`.  Processes in chunks so scratch stays bounded.
int exact_classify(Ctx& c, const int64_t* list, int64_t base, int64_t n, uint8_t* sign, int64_t* fp64_total) {
    const int prec = c.p->winding_precision;
    const int64_t chunk = std::min<int64_t>(n, 8ll << 20);
    for (int64_t off = 0; off < n; off += chunk) {
        int64_t m = std::min(chunk, n - off);
        double* w = (double*)ws_get(c.w, S_W, sizeof(double) * (size_t)m);
        void* scr = ws_get(c.w, S_WSCR, winding_scratch_bytes(m, c.nt));
        int64_t* f64 = (int64_t*)ws_get(c.w, S_FP64, sizeof(int64_t) * (size_t)m);
        if (!w || !scr || !f64) return set_err(RHZ_ERR_NOMEM, "winding scratch allocation failed");
        const int64_t* idx = list ? list + off : nullptr;
        int64_t b = list ? 0 : base + off;
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_FP64), 0, sizeof(unsigned long long), c.s));
        if (c.wind_t) c.wind_t->begin(c.s);
        HIP_TRY(launch_winding(c.tris, c.nt, c.g, idx, b, m, prec, c.p->fp64_band, w, scr, f64,
                               cnt_ptr(c.w, C_FP64), c.s));
        if (c.wind_t) c.wind_t->end(c.s);
        c.st->winding_pairs += m * c.nt;
        if (prec != 0) {
            int rc = read_counters(c);
            if (rc) return rc;
            int64_t nf = (int64_t)hc(c, C_FP64);
            if (fp64_total) *fp64_total += nf;
            HIP_TRY(launch_winding_fp64(c.tris, c.nt, c.g, idx, b, f64, nf, w, c.s));
        }
        if (idx) {
            HIP_TRY(launch_scatter_decision(w, idx, m, sign, c.s));
        } else {
            HIP_TRY(launch_range_decision(w, b, m, sign, c.s));
        }
    }
    return RHZ_OK;
}

// [C] exact fixup of every sign == 2 voxel; returns the count in *fixed.
int fixup_unknown(Ctx& c, uint8_t* sign, int64_t* fixed) {
    const int64_t V = c.g.n();
    int64_t cap = (int64_t)(c.w->bytes[S_LIST] / sizeof(int64_t));
    if (cap < (1 << 20)) cap = 1 << 20;
    for (int attempt = 0; attempt < 2; ++attempt) {
        int64_t* list = (int64_t*)ws_get(c.w, S_LIST, sizeof(int64_t) * (size_t)cap);
        if (!list) return set_err(RHZ_ERR_NOMEM, "unknown-voxel list allocation failed");
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_LIST), 0, sizeof(unsigned long long), c.s));
        HIP_TRY(launch_compact_unknown(sign, V, list, cap, cnt_ptr(c.w, C_LIST), c.s));
        int rc = read_counters(c);
        if (rc) return rc;
        int64_t n = (int64_t)hc(c, C_LIST);
        if (n <= cap) {
            *fixed = n;
            if (n == 0) return RHZ_OK;
            return exact_classify(c, list, 0, n, sign, &c.st->fp64_voxels);
        }
        cap = n;
    }
    return set_err(RHZ_ERR_HIP, "unknown-voxel compaction did not converge");
}

// One round on the lane kernels: x+, x-, y+, y-, z+, z- (cuda_rasterizer.cu
// :1012-1017).  `in` (x-fastest, layout lin) is transposed into b (y-fastest,
// layout ly); x+ b -> a and x- a -> b on that copy; b is transposed back into
// a (x-fastest, layout lx); y and z alternate a and b, so the round ends in
// a.  `in` may be a itself.  Every sweep launch is bracketed by its own
// events; the transposes are timed apart.
int sweep_round_p(Ctx& c, const float* in, const FieldLayout& lin, float* a, float* b, const FieldLayout& lx,
                  const FieldLayout& ly, int* changed) {
    auto sw0 = [&]() { if (c.sweeps_t) c.sweeps_t->begin(c.s); };
    auto sw1 = [&]() { if (c.sweeps_t) c.sweeps_t->end(c.s); };
    auto tr0 = [&]() { if (c.trans_t) c.trans_t->begin(c.s); };
    auto tr1 = [&]() { if (c.trans_t) c.trans_t->end(c.s); };
    tr0();
    HIP_TRY(launch_transpose_xy(in, b, c.g, true, c.s, lin, ly));
    tr1();
    sw0();
    HIP_TRY(launch_sweep_yx(b, a, c.g, +1, changed, c.s, ly));
    sw1();
    sw0();
    HIP_TRY(launch_sweep_yx(a, b, c.g, -1, changed, c.s, ly));
    sw1();
    tr0();
    HIP_TRY(launch_transpose_xy(b, a, c.g, false, c.s, ly, lx));
    tr1();
    const int axes[4] = {1, 1, 2, 2};
    const int dirs[4] = {+1, -1, +1, -1};
    float* src = a;
    float* dst = b;
    for (int k = 0; k < 4; ++k) {
        sw0();
        HIP_TRY(launch_sweep(src, dst, c.g, axes[k], dirs[k], changed, c.s, lx));
        sw1();
        std::swap(src, dst);
    }
    c.st->sweeps += 6;
    return RHZ_OK;
}

// One round with every sweep on the y-fastest copy: the field (in, layout lin)
// is transposed into a (layout ly, rows padded to 32 floats) unless it is
// there already (in_y); then x+, x-, y+, y-, z+, z- alternate a and b and the
// round ends in a, still y-fastest.  No transposes between rounds.
int sweep_round_y(Ctx& c, const float* in, const FieldLayout& lin, bool in_y, float* a, float* b,
                  const FieldLayout& ly, int* changed) {
    if (!in_y) {
        if (c.trans_t) c.trans_t->begin(c.s);
        HIP_TRY(launch_transpose_xy(in, a, c.g, true, c.s, lin, ly));
        if (c.trans_t) c.trans_t->end(c.s);
    }
    const int axes[6] = {0, 0, 1, 1, 2, 2};
    const int dirs[6] = {+1, -1, +1, -1, +1, -1};
    float* src = a;
    float* dst = b;
    for (int k = 0; k < 6; ++k) {
        if (c.sweeps_t) c.sweeps_t->begin(c.s);
        HIP_TRY(launch_sweep_on_yx(src, dst, c.g, axes[k], dirs[k], changed, c.s, ly));
        if (c.sweeps_t) c.sweeps_t->end(c.s);
        std::swap(src, dst);
    }
    c.st->sweeps += 6;
    return RHZ_OK;
}

// One round on a dense field in a (b: the other buffer); the round ends in a.
int sweep_round(Ctx& c, float* a, float* b, int* changed) {
    if (lane_sweeps_ok(c.g)) return sweep_round_p(c, a, dense_xy(c.g), a, b, dense_xy(c.g), dense_yx(c.g), changed);
    const int axes[6] = {0, 0, 1, 1, 2, 2};
    const int dirs[6] = {+1, -1, +1, -1, +1, -1};
    float* src = a;
    float* dst = b;
    for (int k = 0; k < 6; ++k) {
        if (c.sweeps_t) c.sweeps_t->begin(c.s);
        HIP_TRY(launch_sweep(src, dst, c.g, axes[k], dirs[k], changed, c.s));
        if (c.sweeps_t) c.sweeps_t->end(c.s);
        std::swap(src, dst);
    }
    c.st->sweeps += 6;
    return RHZ_OK;
}

// One round of the 8 classical orderings (sz outermost, +1 first, then sy,
// then sx) in place on the dense field c.d.
int octant_round(Ctx& c, int* changed) {
    for (int sz = 1; sz >= -1; sz -= 2)
        for (int sy = 1; sy >= -1; sy -= 2)
            for (int sx = 1; sx >= -1; sx -= 2) {
                if (c.sweeps_t) c.sweeps_t->begin(c.s);
                HIP_TRY(launch_sweep_octant(c.d, c.g, sx, sy, sz, changed, c.s));
                if (c.sweeps_t) c.sweeps_t->end(c.s);
            }
    c.st->sweeps += 8;
    return RHZ_OK;
}

// Octant rounds until one changes nothing (at most cap), each a replay of one
// captured round.  The graph is recorded on c.s without per-sweep events;
// ms_sweeps then counts each replay as one interval.
int octant_rounds_graph(Ctx& c, int cap, int* changed, int* rounds_out) {
    EventAcc* sw = c.sweeps_t;
    hipGraph_t gr = nullptr;
    hipGraphExec_t ge = nullptr;
    HIP_TRY(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
    c.sweeps_t = nullptr;  // no event records inside the capture
    int rc = octant_round(c, changed);
    c.st->sweeps -= 8;  // counted per replay below
    const hipError_t ec = hipStreamEndCapture(c.s, &gr);
    c.sweeps_t = sw;
    hipError_t e = ec;
    if (rc == RHZ_OK && e == hipSuccess) e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    for (int r = 0; rc == RHZ_OK && e == hipSuccess && r < cap; ++r) {
        e = launch_zero_i32(changed, 1, c.s);
        if (e != hipSuccess) break;
        if (sw) sw->begin(c.s);
        e = hipGraphLaunch(ge, c.s);
        if (sw) sw->end(c.s);
        if (e != hipSuccess) break;
        c.st->sweeps += 8;
        rc = read_counters(c);
        *rounds_out = r + 1;
        if (rc == RHZ_OK && hf(c, F_CHANGED) == 0) break;
    }
    if (ge) (void)hipGraphExecDestroy(ge);
    if (gr) (void)hipGraphDestroy(gr);
    if (rc) return rc;
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_err(RHZ_ERR_HIP, "octant rounds: %s", hipGetErrorString(e));
    }
    return RHZ_OK;
}

// Sentinel count of the x-fastest field d in layout L (the caller's dense
// field when d is null).
int audit(Ctx& c, int64_t* far, const float* d = nullptr, const FieldLayout* L = nullptr) {
    HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_FAR), 0, sizeof(unsigned long long), c.s));
    const FieldLayout dense = dense_xy(c.g);
    if (!d || !L || (L->pitch == dense.pitch && L->plane == dense.plane))
        HIP_TRY(launch_count_far(d ? d : c.d, c.g.n(), cnt_ptr(c.w, C_FAR), c.s));
    else
        HIP_TRY(launch_count_far_rows(d, L->pitch, L->plane, c.g.dims[0], c.g.dims[1], c.g.dims[2],
                                      cnt_ptr(c.w, C_FAR), c.s));
    int rc = read_counters(c);
    if (rc) return rc;
    *far = (int64_t)hc(c, C_FAR);
    return RHZ_OK;
}

// All `cap` rounds enqueued without a host round trip in between, each round
// raising its own flag; optionally captured and replayed as one hipGraph.
// The policy's round count is the first round that changed nothing (+1): the
// rounds after it see an unchanged field and change nothing either, so the
// result equals the round-by-round loop.
template <typename Round>
int fsm_rounds_batched(Ctx& c, int cap, int* rounds_out, Round&& round) {
    int* flags = flag_ptr(c.w, F_ROUND0);
    const bool graph = sweep_tuning().fsm_graph != 0 && c.s != nullptr;
    EventAcc* sw = c.sweeps_t;
    EventAcc* tr = c.trans_t;
    hipGraph_t gr = nullptr;
    hipGraphExec_t ge = nullptr;
    if (graph) {
        c.sweeps_t = nullptr;  // no event records inside the capture
        c.trans_t = nullptr;
        HIP_TRY(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
    }
    int rc = RHZ_OK;
    hipError_t e = launch_zero_i32(flags, cap, c.s);
    if (e != hipSuccess) rc = set_err(RHZ_ERR_HIP, "flag reset failed: %s", hipGetErrorString(e));
    for (int r = 0; r < cap && rc == RHZ_OK; ++r) rc = round(flags + r);
    if (graph) {
        const hipError_t ec = hipStreamEndCapture(c.s, &gr);
        c.sweeps_t = sw;
        c.trans_t = tr;
        if (rc == RHZ_OK && ec != hipSuccess) rc = set_err(RHZ_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
        if (rc == RHZ_OK) {
            e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
            if (e == hipSuccess) e = hipGraphLaunch(ge, c.s);
            if (e != hipSuccess) rc = set_err(RHZ_ERR_HIP, "graph launch failed: %s", hipGetErrorString(e));
        }
    }
    if (rc == RHZ_OK) rc = read_counters(c);  // synchronises
    if (ge) (void)hipGraphExecDestroy(ge);
    if (gr) (void)hipGraphDestroy(gr);
    if (rc) return rc;
    int rounds = cap;
    for (int r = 0; r < cap; ++r)
        if (hf(c, F_ROUND0 + r) == 0) {
            rounds = r + 1;
            break;
        }
    *rounds_out = rounds;
    return RHZ_OK;
}

// Where the FSM rounds left the field: the caller's dense buffer, or a padded
// working copy in layout l (y fastest when yfast).  deferred: the [D2] audit
// and the extension rounds are left to run_fsm_finish, which audits while it
// writes the copy back; `other` is the second working buffer for those rounds.
struct FsmField {
    const float* d;
    FieldLayout l;
    bool yfast;
    bool deferred = false;
    float* other = nullptr;
};

// [D2] audit of a y-fastest copy d (rows along y, pitch and plane of L).
int audit_rows_y(Ctx& c, int64_t* far, const float* d, const FieldLayout& L) {
    HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_FAR), 0, sizeof(unsigned long long), c.s));
    HIP_TRY(launch_count_far_rows(d, L.pitch, L.plane, c.g.dims[1], c.g.dims[0], c.g.dims[2], cnt_ptr(c.w, C_FAR),
                                  c.s));
    const int rc = read_counters(c);
    if (rc) return rc;
    *far = (int64_t)hc(c, C_FAR);
    return RHZ_OK;
}

// [D] far-field re-distancing with the reference round policy: the rounds
// and the audit-gated extension rounds; *far = voxels still at the sentinel.
// The rounds run on two working buffers with padded rows (fsm_layouts); the
// first round reads the caller's dense field, and the result is copied back
// into it at the end, or, when `keep` is given, left where it is and named in
// *keep (run_fsm_finish writes it back).  Lattices the lane kernels do not
// take (or that are already aligned) run in place on the dense field.
int run_fsm_rounds(Ctx& c, int64_t* far_out, FsmField* keep = nullptr) {
    const int64_t V = c.g.n();
    const int64_t dx = c.g.dims[0], dz = c.g.dims[2];
    // RHZ_FSM_OCTANT: rounds of the 8 classical orderings in place on the
    // dense field until nothing changes
    const bool octant = c.p->fsm_policy == RHZ_FSM_OCTANT;
    const bool lanes = !octant && lane_sweeps_ok(c.g);
    const FieldLayout dense = dense_xy(c.g);
    FieldLayout lx = dense, ly = dense_yx(c.g);
    if (lanes) fsm_layouts(c.g, &lx, &ly);
    const bool padded = lx.plane != dense.plane || ly.plane != dense.plane;
    const size_t cells = padded ? (size_t)(std::max(lx.plane, ly.plane) * dz) : (size_t)V;
    float* q = octant ? nullptr : (float*)ws_get(c.w, S_DIST2, sizeof(float) * cells);
    float* pbuf = padded ? (float*)ws_get(c.w, S_DIST3, sizeof(float) * cells) : c.d;
    if ((!octant && !q) || !pbuf) return set_err(RHZ_ERR_NOMEM, "sweep buffer allocation failed");
    const float* cur = c.d;  // the field: the caller's dense buffer until a round has run
    FieldLayout cur_l = dense;
    // every sweep on the y-fastest copy (tuning "fsm_layout" 1): the field stays
    // y-fastest from the first round to the end of the rounds
    const bool ylay = lanes && padded && sweep_tuning().fsm_layout == 1 && ly.pitch % 32 == 0;
    bool cur_y = false;
    auto round = [&](int* flag) -> int {
        if (octant) return octant_round(c, flag);
        if (!lanes) return sweep_round(c, c.d, q, flag);
        if (ylay) {
            const int rc = sweep_round_y(c, cur, cur_l, cur_y, pbuf, q, ly, flag);
            cur = pbuf;
            cur_l = ly;
            cur_y = true;
            return rc;
        }
        const int rc = sweep_round_p(c, cur, cur_l, pbuf, q, lx, ly, flag);
        cur = pbuf;
        cur_l = lx;
        return rc;
    };
    // sentinel count of the current field (rows along y on the y-fastest copy)
    auto audit_cur = [&](int64_t* far) -> int {
        if (!cur_y) return audit(c, far, cur, &cur_l);
        return audit_rows_y(c, far, cur, cur_l);
    };
    int* changed = flag_ptr(c.w, F_CHANGED);
    const bool converge = c.p->fsm_policy == RHZ_FSM_CONVERGE || octant;
    const int cap = converge ? c.p->fsm_round_cap : c.p->max_rounds;
    int rounds = 0;
    const SweepTuning tune = sweep_tuning();
    if (octant && cap > 0 && c.s != nullptr) {
        // one round (8 orderings, a launch per diagonal level) captured once
        // and replayed per round: the flag is cleared by a kernel before each
        // replay and read after it
        int rc = octant_rounds_graph(c, cap, changed, &rounds);
        if (rc) return rc;
    } else if (!converge && cap > 0 && cap <= kMaxBatchedRounds && V <= tune.batch_rounds_max_voxels) {
        int rc = fsm_rounds_batched(c, cap, &rounds, round);
        if (rc) return rc;
    } else
    for (int r = 0; r < cap; ++r) {
        HIP_TRY(hipMemsetAsync(changed, 0, sizeof(int), c.s));
        int rc = round(changed);
        if (rc) return rc;
        rc = read_counters(c);
        if (rc) return rc;
        rounds = r + 1;
        if (hf(c, F_CHANGED) == 0) break;
    }
    int64_t far = 0;
    int extra = 0;
    // A y-fastest copy the caller keeps is audited by its write-back, which
    // counts the sentinel cells as it transposes them (run_fsm_finish runs
    // the extension rounds there when any are left): one pass over the field
    // fewer.
    const bool defer = cur_y && keep;
    if (!defer) {
        int rc = audit_cur(&far);
        if (rc) return rc;
        while (far > 0 && extra < c.p->max_extra_rounds) {
            rc = round(changed);
            if (rc) return rc;
            ++extra;
            rc = audit_cur(&far);
            if (rc) return rc;
        }
    }
    if (defer) {
        *keep = FsmField{cur, ly, true, true, q};  // run_fsm_finish audits and transposes it back
    } else if (cur_y) {
        // back to the x-fastest layout, straight into the caller's field
        if (c.trans_t) c.trans_t->begin(c.s);
        HIP_TRY(launch_transpose_xy(cur, c.d, c.g, false, c.s, ly, dense));
        if (c.trans_t) c.trans_t->end(c.s);
    } else if (keep) {
        *keep = FsmField{cur, cur_l, false};
    } else if (cur != c.d) {
        HIP_TRY(launch_copy_rows(cur, cur_l.pitch, cur_l.plane, c.d, dense.pitch, dense.plane, (int)dx, c.g.dims[1],
                                 (int)dz, c.s));
    }
    c.st->eikonal_rounds = rounds + extra;
    c.st->extra_rounds = extra;
    *far_out = defer ? -1 : far;
    return RHZ_OK;
}

// [D3] safe clamp of what the rounds left at the sentinel (needs the signs).
int run_fsm_clamp(Ctx& c, const uint8_t* sign, int64_t far) {
    const int64_t V = c.g.n();
    c.st->safe_clamped = 0;
    if (far > 0) {
        c.st->safe_clamped = far;
        HIP_TRY(launch_clamp_far(c.d, sign, V, (float)c.p->band_voxels * c.g.h, c.p->clamp_intended_sign, c.s));
        const int rc = audit(c, &far);
        if (rc) return rc;
    }
    c.st->far_residual = far;
    return RHZ_OK;
}

int run_fsm(Ctx& c, const uint8_t* sign) {
    int64_t far = 0;
    const int rc = run_fsm_rounds(c, &far);
    if (rc) return rc;
    return run_fsm_clamp(c, sign, far);
}

// [D3] and [F] after run_fsm_rounds kept its result in field f.  With voxels
// left at the sentinel, f goes back into the dense field and the clamp and the
// apply run there; otherwise (the clamp has nothing to do) the write-back
// applies the signs on the way: one pass over the field instead of two (for a
// y-fastest copy, the transpose back to x fastest is that pass).
// mark() runs on the stream just before the apply.
template <typename Mark>
int run_fsm_finish(Ctx& c, const uint8_t* sign, int64_t far, const FsmField& f, Mark&& mark) {
    const FieldLayout dense = dense_xy(c.g);
    if (f.deferred) {
        // the write-back with the signs, auditing the copy on the way; when
        // nothing is left at the sentinel (the usual case) it is final
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_FAR), 0, sizeof(unsigned long long), c.s));
        mark();
        HIP_TRY(launch_transpose_apply(f.d, f.l, c.d, sign, c.g, c.s, cnt_ptr(c.w, C_FAR)));
        int rc = read_counters(c);
        if (rc) return rc;
        far = (int64_t)hc(c, C_FAR);
        if (far == 0) return run_fsm_clamp(c, sign, 0);
        // cells left at the sentinel: the extension rounds on the copy (still
        // intact), then the write-back without the fusion over what was written
        int* changed = flag_ptr(c.w, F_CHANGED);
        int extra = 0;
        while (far > 0 && extra < c.p->max_extra_rounds) {
            rc = sweep_round_y(c, f.d, f.l, true, const_cast<float*>(f.d), f.other, f.l, changed);
            if (rc) return rc;
            ++extra;
            rc = audit_rows_y(c, &far, f.d, f.l);
            if (rc) return rc;
        }
        c.st->eikonal_rounds += extra;
        c.st->extra_rounds = extra;
        HIP_TRY(launch_transpose_xy(f.d, c.d, c.g, false, c.s, f.l, dense));
        rc = run_fsm_clamp(c, sign, far);
        if (rc) return rc;
        mark();
        HIP_TRY(launch_apply(c.d, sign, c.g.n(), c.s));
        return RHZ_OK;
    }
    const bool fused = f.d != c.d && far == 0;
    if (f.d != c.d && !fused) {
        if (f.yfast)
            HIP_TRY(launch_transpose_xy(f.d, c.d, c.g, false, c.s, f.l, dense));
        else
            HIP_TRY(launch_copy_rows(f.d, f.l.pitch, f.l.plane, c.d, dense.pitch, dense.plane, c.g.dims[0],
                                     c.g.dims[1], c.g.dims[2], c.s));
    }
    const int rc = run_fsm_clamp(c, sign, far);
    if (rc) return rc;
    mark();
    if (fused && f.yfast)
        HIP_TRY(launch_transpose_apply(f.d, f.l, c.d, sign, c.g, c.s));
    else if (fused)
        HIP_TRY(launch_copy_apply_rows(f.d, f.l.pitch, f.l.plane, c.d, sign, c.g.dims[0], c.g.dims[1], c.g.dims[2],
                                       c.s));
    else
        HIP_TRY(launch_apply(c.d, sign, c.g.n(), c.s));
    return RHZ_OK;
}

bool int_range_ok(const Grid& g, int64_t nt) {
    const int64_t dx = g.dims[0], dy = g.dims[1], dz = g.dims[2];
    // cuda_rasterizer.cu:815-820
    if (g.n() > 2000000000LL) return false;
    if (nt * dz > 2000000000LL) return false;
    if (dx * dy > 2000000000LL || dx * dz > 2000000000LL || dy * dz > 2000000000LL) return false;
    return true;
}

// [B] candidate lists + parity (or the signed crossing count); returns
// RHZ_DECLINED on a pathological list.  *csr receives the lists.
int run_parity(Ctx& c, uint8_t* sign, int* unknown_flag, bool signed_count = false, CsrBuffers* csr = nullptr) {
    const rhz_sdf_params* p = c.p;
    const int sup = p->sup;
    const int64_t nsup = num_supercells(c.g, sup);
    CsrBuffers b;
    b.xyab = (float4*)ws_get(c.w, S_XYAB, sizeof(float4) * (size_t)c.nt);
    b.cover = (int4*)ws_get(c.w, S_COVER, sizeof(int4) * (size_t)c.nt);
    b.counts = (uint32_t*)ws_get(c.w, S_COUNTS, sizeof(uint32_t) * (size_t)nsup);
    b.start = (int64_t*)ws_get(c.w, S_START, sizeof(int64_t) * (size_t)(nsup + 1));
    b.cursor = (uint32_t*)ws_get(c.w, S_CURSOR, sizeof(uint32_t) * (size_t)nsup);
    void* scan = ws_get(c.w, S_SCAN, scan_scratch_bytes(nsup));
    if (!b.xyab || !b.cover || !b.counts || !b.start || !b.cursor || !scan)
        return set_err(RHZ_ERR_NOMEM, "candidate-list allocation failed");
    HIP_TRY(launch_csr_count(c.tris, c.nt, c.g, sup, b, cnt_ptr(c.w, C_CSR), c.s));
    int rc = read_counters(c);
    if (rc) return rc;
    int64_t entries = (int64_t)hc(c, C_CSR);
    c.st->csr_entries = entries;
    if (entries <= 0 || entries > p->csr_max_entries) {
        c.st->decline_reason = 2;
        return RHZ_DECLINED;  // pathological triangle spread (:867-869)
    }
    b.ids = (int32_t*)ws_get(c.w, S_IDS, sizeof(int32_t) * (size_t)entries);
    if (!b.ids) return set_err(RHZ_ERR_NOMEM, "candidate-list allocation failed");
    HIP_TRY(launch_scan_u32(b.counts, nsup, b.start, scan, c.s));
    HIP_TRY(launch_csr_fill(c.nt, c.g, sup, b, c.s));
    HIP_TRY(launch_parity(c.tris, c.g, sup, b, p->jitter_x, p->jitter_y, p->parity_max_hits, sign,
                          flag_ptr(c.w, F_UNKNOWN), c.s, signed_count));
    rc = read_counters(c);
    if (rc) return rc;
    *unknown_flag = hf(c, F_UNKNOWN);
    if (csr) *csr = b;
    return RHZ_OK;
}

// [E] stratified exact verification with patch boxes.  Returns RHZ_OK when a
// probe comes back clean, RHZ_DECLINED otherwise.
int run_verify(Ctx& c, uint8_t* sign) {
    const rhz_sdf_params* p = c.p;
    const int N = p->verify_samples;
    const int64_t V = c.g.n();
    int64_t* idx = (int64_t*)ws_get(c.w, S_SAMPLES, sizeof(int64_t) * (size_t)N);
    int64_t* sites = (int64_t*)ws_get(c.w, S_SITES, sizeof(int64_t) * (size_t)N);
    double* w = (double*)ws_get(c.w, S_MISC, sizeof(double) * (size_t)N);
    void* scr = ws_get(c.w, S_WSCR, winding_scratch_bytes(N, c.nt));
    int64_t* f64 = (int64_t*)ws_get(c.w, S_FP64, sizeof(int64_t) * (size_t)N);
    if (!idx || !sites || !w || !scr || !f64) return set_err(RHZ_ERR_NOMEM, "verify scratch allocation failed");
    bool ok = false;
    for (int gen = 0; gen < p->verify_gens; ++gen) {
        HIP_TRY(launch_verify_indices(V, N, gen, idx, c.s));
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_FP64), 0, 2 * sizeof(unsigned long long), c.s));
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_MISM), 0, sizeof(unsigned long long), c.s));
        if (c.wind_t) c.wind_t->begin(c.s);
        HIP_TRY(launch_winding(c.tris, c.nt, c.g, idx, 0, N, p->winding_precision, p->fp64_band, w, scr, f64,
                               cnt_ptr(c.w, C_FP64), c.s));
        if (c.wind_t) c.wind_t->end(c.s);
        c.st->winding_pairs += (int64_t)N * c.nt;
        if (p->winding_precision != 0) {
            int rc = read_counters(c);
            if (rc) return rc;
            int64_t nf = (int64_t)hc(c, C_FP64);
            c.st->fp64_voxels += nf;
            HIP_TRY(launch_winding_fp64(c.tris, c.nt, c.g, idx, 0, f64, nf, w, c.s));
        }
        HIP_TRY(launch_compare_samples(w, idx, N, sign, sites, N, cnt_ptr(c.w, C_MISM), c.s));
        int rc = read_counters(c);
        if (rc) return rc;
        int m = (int)std::min<unsigned long long>(hc(c, C_MISM), (unsigned long long)N);
        c.st->probes++;
        c.st->last_probe_m = m;
        if (m == 0) {
            ok = true;
            break;
        }
        if (m > p->max_sites_per_probe) break;  // systematic divergence
        if (gen == p->verify_gens - 1) break;   // persistent after the patches
        HIP_TRY(launch_mark_patch(sign, sites, m, p->patch_radius, c.g.dims, c.s));
        int64_t fixed = 0;
        rc = fixup_unknown(c, sign, &fixed);
        if (rc) return rc;
        c.st->patched_voxels += fixed;
        c.st->sites += m;
        c.st->patch_rounds++;
    }
    if (!ok) {
        c.st->decline_reason = 3;
        return RHZ_DECLINED;
    }
    return RHZ_OK;
}

// Reference fast path on device buffers.
int run_fast(Ctx& c) {
    const rhz_sdf_params* p = c.p;
    rhz_sdf_stats* st = c.st;
    const int64_t V = c.g.n();
    if (!int_range_ok(c.g, c.nt) || c.nt <= 0) {
        st->decline_reason = 1;
        return RHZ_DECLINED;
    }
    uint8_t* sign = (uint8_t*)ws_get(c.w, S_SIGN, (size_t)V);
    if (!sign) {
        st->decline_reason = 4;
        return RHZ_DECLINED;  // VRAM allocation failed (:914-918)
    }
    Timer tm(true);
    tm.mark(0, c.s);
    int rc = zero_counters(c);
    if (rc) return rc;
    HIP_TRY(launch_fill(c.d, V, kFar, c.s));
    HIP_TRY(launch_splat(c.tris, c.nt, c.g, (float)p->band_voxels * c.g.h, c.d, c.s));
    tm.mark(1, c.s);
    int unknown = 0;
    rc = run_parity(c, sign, &unknown);
    if (rc) return rc;
    tm.mark(2, c.s);
    st->unknown_flag = unknown;
    if (unknown) {
        int64_t fixed = 0;
        rc = fixup_unknown(c, sign, &fixed);
        if (rc) return rc;
        st->fixup_voxels = fixed;
    }
    tm.mark(3, c.s);
    rc = run_fsm(c, sign);
    if (rc) return rc;
    tm.mark(4, c.s);
    rc = run_verify(c, sign);
    if (rc) return rc;
    tm.mark(5, c.s);
    HIP_TRY(launch_apply(c.d, sign, V, c.s));
    tm.mark(6, c.s);
    HIP_TRY(hipStreamSynchronize(c.s));
    st->ms_splat = tm.ms(0, 1);
    st->ms_parity = tm.ms(1, 2);
    st->ms_fixup = tm.ms(2, 3);
    st->ms_eikonal = tm.ms(3, 4);
    st->ms_verify = tm.ms(4, 5);
    st->ms_apply = tm.ms(5, 6);
    st->ms_total = tm.ms(0, 6);
    return RHZ_OK;
}

int run_legacy(Ctx& c) {
    void* bb = ws_get(c.w, S_MISC, legacy_scratch_bytes(c.nt));
    if (!bb) return set_err(RHZ_ERR_NOMEM, "legacy scratch allocation failed");
    Timer tm(true);
    tm.mark(0, c.s);
    HIP_TRY(launch_legacy(c.tris, c.nt, c.g, c.d, bb, c.s));
    tm.mark(1, c.s);
    HIP_TRY(hipStreamSynchronize(c.s));
    c.st->ms_total = tm.ms(0, 1);
    return RHZ_OK;
}

// Exact winding sign for every voxel, exact band, FSM, apply.
int run_plain_exact(Ctx& c) {
    const rhz_sdf_params* p = c.p;
    const int64_t V = c.g.n();
    uint8_t* sign = (uint8_t*)ws_get(c.w, S_SIGN, (size_t)V);
    if (!sign) return set_err(RHZ_ERR_NOMEM, "sign volume allocation failed");
    Timer tm(true);
    tm.mark(0, c.s);
    int rc = zero_counters(c);
    if (rc) return rc;
    HIP_TRY(launch_fill(c.d, V, kFar, c.s));
    HIP_TRY(launch_splat(c.tris, c.nt, c.g, (float)p->band_voxels * c.g.h, c.d, c.s));
    tm.mark(1, c.s);
    rc = exact_classify(c, nullptr, 0, V, sign, &c.st->fp64_voxels);
    if (rc) return rc;
    c.st->reprobed_voxels = V;
    tm.mark(2, c.s);
    tm.mark(3, c.s);
    FsmField fld{};
    int64_t far = 0;
    rc = run_fsm_rounds(c, &far, &fld);
    if (rc) return rc;
    rc = run_fsm_finish(c, sign, far, fld, [&] {
        tm.mark(4, c.s);
        tm.mark(5, c.s);
    });
    if (rc) return rc;
    tm.mark(6, c.s);
    HIP_TRY(hipStreamSynchronize(c.s));
    c.st->ms_splat = tm.ms(0, 1);
    c.st->ms_parity = tm.ms(1, 2);
    c.st->ms_eikonal = tm.ms(3, 4);
    c.st->ms_apply = tm.ms(5, 6);
    c.st->ms_total = tm.ms(0, 6);
    return RHZ_OK;
}

// ---------------------------------------------------------------------------
// learned sign gate (FAST_LEARNED)
// ---------------------------------------------------------------------------
#if __has_include("proposer_weights.inc")
#include "proposer_weights.inc"  // kProposerBlob[9241]: weights trained by the project's own kernels
#define RHZ_HAVE_BUILTIN_PROPOSER 1
#endif

// Spatial clusters of the call's triangles for the far-field certificate:
// cells of a grid of side sqrt(24 x mean triangle area) over the triangles'
// box (at most 4M cells), by centroid.
int build_clusters(Ctx& c) {
    if (c.ncl >= 0) return RHZ_OK;
    const int nblocks = (int)std::min<int64_t>(1024, (c.nt + 255) / 256);
    char* boxbuf = (char*)ws_get(c.w, S_CL_BOX, 64 + sizeof(double) * (size_t)nblocks);
    if (!boxbuf) return set_err(RHZ_ERR_NOMEM, "cluster scratch allocation failed");
    int* box = reinterpret_cast<int*>(boxbuf);
    double* area_part = reinterpret_cast<double*>(boxbuf + 64);
    const int init[6] = {0x7fffffff, 0x7fffffff, 0x7fffffff, (int)0x80000000, (int)0x80000000, (int)0x80000000};
    HIP_TRY(hipMemcpyAsync(box, init, sizeof(init), hipMemcpyHostToDevice, c.s));
    HIP_TRY(launch_tri_box_area(c.tris, c.nt, box, area_part, nblocks, c.s));
    std::vector<char> hb(64 + sizeof(double) * (size_t)nblocks);
    HIP_TRY(hipMemcpyAsync(hb.data(), boxbuf, hb.size(), hipMemcpyDeviceToHost, c.s));
    HIP_TRY(hipStreamSynchronize(c.s));
    const int* hbox = reinterpret_cast<const int*>(hb.data());
    const double* hpart = reinterpret_cast<const double*>(hb.data() + 64);
    auto unorder = [](int i) {
        const int b = i >= 0 ? i : i ^ 0x7fffffff;
        float f;
        std::memcpy(&f, &b, 4);
        return f;
    };
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = unorder(hbox[k]);
        hi[k] = unorder(hbox[3 + k]);
    }
    double area = 0.0;
    for (int b = 0; b < nblocks; ++b) area += hpart[b];
    double side = std::sqrt(24.0 * area / (double)std::max<int64_t>(c.nt, 1));
    double ext = 0.0;
    for (int k = 0; k < 3; ++k) ext = std::max(ext, (double)hi[k] - (double)lo[k]);
    if (!(side > 0.0) || !std::isfinite(side)) side = std::max(ext, 1e-3);
    side = std::max(side, ext / 1024.0);
    ClusterGrid g;
    int n3[3];
    for (;;) {
        int64_t cells = 1;
        for (int k = 0; k < 3; ++k) {
            n3[k] = (int)std::floor(((double)hi[k] - (double)lo[k]) / side) + 1;
            cells *= n3[k];
        }
        if (cells <= (4ll << 20)) break;
        side *= 1.26;
    }
    g.ox = lo[0];
    g.oy = lo[1];
    g.oz = lo[2];
    g.inv_s = (float)(1.0 / side);
    g.nx = n3[0];
    g.ny = n3[1];
    g.nz = n3[2];
    const int64_t ncells = g.cells(), npar = g.parents();
    uint32_t* counts = (uint32_t*)ws_get(c.w, S_CL_COUNTS, sizeof(uint32_t) * (size_t)ncells);
    uint32_t* cursor = (uint32_t*)ws_get(c.w, S_CL_CURSOR, sizeof(uint32_t) * (size_t)ncells);
    int64_t* start = (int64_t*)ws_get(c.w, S_CL_START, sizeof(int64_t) * (size_t)(ncells + 1));
    void* scan = ws_get(c.w, S_CL_SCAN, scan_scratch_bytes(ncells));
    int32_t* ids = (int32_t*)ws_get(c.w, S_CL_IDS, sizeof(int32_t) * (size_t)c.nt);
    float4* info = (float4*)ws_get(c.w, S_CL_INFO, sizeof(float4) * 5 * (size_t)ncells);
    int2* range = (int2*)ws_get(c.w, S_CL_RANGE, sizeof(int2) * (size_t)ncells);
    int64_t* slot = (int64_t*)ws_get(c.w, S_CL_SLOT, sizeof(int64_t) * (size_t)(ncells + 1));
    uint32_t* pflag = (uint32_t*)ws_get(c.w, S_CL_PFLAG, sizeof(uint32_t) * (size_t)npar);
    int64_t* pslot = (int64_t*)ws_get(c.w, S_CL_PSLOT, sizeof(int64_t) * (size_t)(npar + 1));
    float4* pinfo = (float4*)ws_get(c.w, S_CL_PINFO, sizeof(float4) * 5 * (size_t)npar);
    int2* prange = (int2*)ws_get(c.w, S_CL_PRANGE, sizeof(int2) * (size_t)npar);
    if (!counts || !cursor || !start || !scan || !ids || !info || !range || !slot || !pflag || !pslot || !pinfo ||
        !prange)
        return set_err(RHZ_ERR_NOMEM, "cluster allocation failed");
    HIP_TRY(hipMemsetAsync(counts, 0, sizeof(uint32_t) * (size_t)ncells, c.s));
    HIP_TRY(hipMemsetAsync(cursor, 0, sizeof(uint32_t) * (size_t)ncells, c.s));
    HIP_TRY(launch_cluster_build(c.tris, c.nt, g, counts, cursor, start, scan, ids, info, range, slot, pflag, pslot,
                                 pinfo, prange, c.s));
    int64_t nums[2];
    HIP_TRY(hipMemcpyAsync(&nums[0], slot + ncells, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
    HIP_TRY(hipMemcpyAsync(&nums[1], pslot + npar, sizeof(int64_t), hipMemcpyDeviceToHost, c.s));
    HIP_TRY(hipStreamSynchronize(c.s));
    c.ncl = (int)nums[0];
    c.npar = (int)nums[1];
    c.st->n_parent_clusters = c.npar;
    c.cl_grid = g;
    c.cl_info = info;
    c.cl_range = range;
    c.cl_pinfo = pinfo;
    c.cl_prange = prange;
    c.cl_ids = ids;
    c.st->n_clusters = c.ncl;
    return RHZ_OK;
}

// Winding numbers of `m` listed voxels into w: far-field certificate at
// beta, brute-force exact kernel for the voxels it leaves undecided.  Probe voxels are taken in spatial order (counting
// sort by cluster cell) so a wave's lanes share their near clusters.
int winding_certified(Ctx& c, const int64_t* list, int64_t m, double* w) {
    int rc = build_clusters(c);
    if (rc) return rc;
    const float beta = std::max(c.p->far_field_beta, 2.0f);  // the bound needs d > r
    const ClusterGrid& cg = c.cl_grid;
    const int64_t ncells = cg.cells();
    uint32_t* counts = (uint32_t*)ws_get(c.w, S_CL_COUNTS, sizeof(uint32_t) * (size_t)ncells);
    uint32_t* cursor = (uint32_t*)ws_get(c.w, S_CL_CURSOR, sizeof(uint32_t) * (size_t)ncells);
    int64_t* start = (int64_t*)ws_get(c.w, S_CL_START, sizeof(int64_t) * (size_t)(ncells + 1));
    void* scan = ws_get(c.w, S_CL_SCAN, scan_scratch_bytes(ncells));
    int64_t* order = (int64_t*)ws_get(c.w, S_ORDER, sizeof(int64_t) * (size_t)m);
    int64_t* und = (int64_t*)ws_get(c.w, S_UNDEC, sizeof(int64_t) * (size_t)m);
    void* part = ws_get(c.w, S_FARPART, far_scratch_bytes(m, c.npar));
    if (!counts || !cursor || !start || !scan || !order || !und || !part)
        return set_err(RHZ_ERR_NOMEM, "certificate scratch allocation failed");
    HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_UNDEC), 0, sizeof(unsigned long long), c.s));
    if (c.far_t) c.far_t->begin(c.s);
    HIP_TRY(launch_point_sort(c.g, list, nullptr, m, cg, counts, cursor, start, scan, order, c.s));
    HIP_TRY(launch_winding_far(c.tris, c.cl_info, c.cl_range, c.cl_ids, c.cl_pinfo, c.cl_prange, c.npar, true, c.g,
                               list, order, m, beta, part, w, und, cnt_ptr(c.w, C_UNDEC), c.s));
    rc = read_counters(c);
    if (rc) return rc;
    int64_t nu = (int64_t)hc(c, C_UNDEC);
    const int64_t* left = und;
    if (nu > 0) {
        // the probes the parents' looser bounds left open, at the child level
        int64_t* und2 = (int64_t*)ws_get(c.w, S_UNDEC2, sizeof(int64_t) * (size_t)nu);
        void* part2 = ws_get(c.w, S_FARPART, far_scratch_bytes(nu, c.npar));
        if (!und2 || !part2) return set_err(RHZ_ERR_NOMEM, "certificate scratch allocation failed");
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_UNDEC2), 0, sizeof(unsigned long long), c.s));
        HIP_TRY(launch_winding_far(c.tris, c.cl_info, c.cl_range, c.cl_ids, c.cl_pinfo, c.cl_prange, c.npar, false,
                                   c.g, list, und, nu, beta, part2, w, und2, cnt_ptr(c.w, C_UNDEC2), c.s));
        rc = read_counters(c);
        if (rc) return rc;
        c.st->far_second_pass += nu;
        nu = (int64_t)hc(c, C_UNDEC2);
        left = und2;
    }
    if (c.far_t) c.far_t->end(c.s);
    c.st->far_decided += m - nu;
    c.st->far_undecided += nu;
    if (nu == 0) return RHZ_OK;
    int64_t* vox = (int64_t*)ws_get(c.w, S_UNDVOX, sizeof(int64_t) * (size_t)nu);
    double* w2 = (double*)ws_get(c.w, S_W2, sizeof(double) * (size_t)nu);
    void* scr = ws_get(c.w, S_WSCR, winding_scratch_bytes(nu, c.nt));
    int64_t* f64 = (int64_t*)ws_get(c.w, S_FP64, sizeof(int64_t) * (size_t)nu);
    if (!vox || !w2 || !scr || !f64) return set_err(RHZ_ERR_NOMEM, "winding scratch allocation failed");
    HIP_TRY(launch_gather_i64(list, left, nu, vox, c.s));
    const int prec = c.p->winding_precision;
    HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_FP64), 0, sizeof(unsigned long long), c.s));
    if (c.wind_t) c.wind_t->begin(c.s);
    HIP_TRY(launch_winding(c.tris, c.nt, c.g, vox, 0, nu, prec, c.p->fp64_band, w2, scr, f64, cnt_ptr(c.w, C_FP64),
                           c.s));
    if (c.wind_t) c.wind_t->end(c.s);
    c.st->winding_pairs += nu * c.nt;
    if (prec != 0) {
        rc = read_counters(c);
        if (rc) return rc;
        const int64_t nf = (int64_t)hc(c, C_FP64);
        c.st->fp64_voxels += nf;
        HIP_TRY(launch_winding_fp64(c.tris, c.nt, c.g, vox, 0, f64, nf, w2, c.s));
    }
    HIP_TRY(launch_scatter_f64(w2, left, nu, w, c.s));
    return RHZ_OK;
}

// Exact winding numbers of `m` listed voxels into w[m]: through the far-field
// certificate (decisions equal to the plain kernel's) or the plain kernel.
int probe_winding(Ctx& c, const int64_t* list, int64_t m, double* w) {
    if (c.p->far_field_beta > 0.0f) return winding_certified(c, list, m, w);
    const int prec = c.p->winding_precision;
    void* scr = ws_get(c.w, S_WSCR, winding_scratch_bytes(m, c.nt));
    int64_t* f64 = (int64_t*)ws_get(c.w, S_FP64, sizeof(int64_t) * (size_t)m);
    if (!scr || !f64) return set_err(RHZ_ERR_NOMEM, "winding scratch allocation failed");
    HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_FP64), 0, sizeof(unsigned long long), c.s));
    if (c.wind_t) c.wind_t->begin(c.s);
    HIP_TRY(launch_winding(c.tris, c.nt, c.g, list, 0, m, prec, c.p->fp64_band, w, scr, f64, cnt_ptr(c.w, C_FP64),
                           c.s));
    if (c.wind_t) c.wind_t->end(c.s);
    c.st->winding_pairs += m * c.nt;
    if (prec != 0) {
        int rc = read_counters(c);
        if (rc) return rc;
        const int64_t nf = (int64_t)hc(c, C_FP64);
        c.st->fp64_voxels += nf;
        HIP_TRY(launch_winding_fp64(c.tris, c.nt, c.g, list, 0, f64, nf, w, c.s));
    }
    return RHZ_OK;
}

// Exact decisions for `n` listed voxels committed into `sign`.  Voxels whose
// sign changed are appended to mism (counter C_MISM) when their list position
// is below n_first, else to mism2 (counter C_MISM2); neither counter is reset
// here.
int probe_commit_ranges(Ctx& c, const int64_t* list, int64_t n, int64_t n_first, uint8_t* sign, int64_t* mism,
                        int64_t mism_cap, int64_t* mism2, int64_t mism2_cap) {
    const int64_t chunk = std::min<int64_t>(n, 8ll << 20);
    for (int64_t off = 0; off < n; off += chunk) {
        const int64_t m = std::min(chunk, n - off);
        double* w = (double*)ws_get(c.w, S_W, sizeof(double) * (size_t)m);
        if (!w) return set_err(RHZ_ERR_NOMEM, "winding buffer allocation failed");
        int rc = probe_winding(c, list + off, m, w);
        if (rc) return rc;
        const int64_t split = std::min(std::max(n_first - off, (int64_t)0), m);
        if (split > 0)
            HIP_TRY(launch_commit(w, list + off, split, sign, mism, mism_cap, cnt_ptr(c.w, C_MISM), c.s));
        if (split < m)
            HIP_TRY(launch_commit(w + split, list + off + split, m - split, sign, mism2, mism2_cap,
                                  cnt_ptr(c.w, C_MISM2), c.s));
    }
    return RHZ_OK;
}

// Exact decisions for `n` listed voxels committed into `sign`; voxels whose
// sign changed are appended to mism (counter C_MISM, not reset here).
int probe_commit(Ctx& c, const int64_t* list, int64_t n, uint8_t* sign, int64_t* mism, int64_t mism_cap) {
    return probe_commit_ranges(c, list, n, n, sign, mism, mism_cap, nullptr, 0);
}

// Band distance (voxels) beyond which a scanline sign cannot be wrong: the
// scanline pass classifies the point (jx, jy) h away from the voxel centre,
// and the two can only differ if the segment between them meets the surface.
float candidate_radius(const rhz_sdf_params* p) {
    if (p->candidate_radius > 0.0f) return p->candidate_radius;
    const double reach = std::sqrt((double)p->jitter_x * p->jitter_x + (double)p->jitter_y * p->jitter_y);
    return (float)(reach * 1.001 + 1e-4);
}

// Candidates + features + certificate flags (after splat, signed scan and
// fixup).  On return *cand / *feats / *risk are workspace buffers holding
// *n_cand rows; *n_risky counts the flagged ones.
// Compacts the candidate voxels (band distance within the candidate radius;
// the last step that reads the band, so it runs before the eikonal rounds).
int learned_candidate_list(Ctx& c, int64_t** cand, int64_t* n_cand) {
    const int64_t V = c.g.n();
    const float thr = candidate_radius(c.p) * c.g.h;
    int64_t cap = (int64_t)(c.w->bytes[S_CAND] / sizeof(int64_t));
    cap = std::max<int64_t>(cap, std::max<int64_t>(1 << 20, V / 512));
    for (int attempt = 0; attempt < 2; ++attempt) {
        int64_t* list = (int64_t*)ws_get(c.w, S_CAND, sizeof(int64_t) * (size_t)cap);
        if (!list) return set_err(RHZ_ERR_NOMEM, "candidate list allocation failed");
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_LIST), 0, sizeof(unsigned long long), c.s));
        HIP_TRY(launch_candidates(c.d, V, thr, list, cap, cnt_ptr(c.w, C_LIST), c.s));
        int rc = read_counters(c);
        if (rc) return rc;
        const int64_t n = (int64_t)hc(c, C_LIST);
        if (n <= cap) {
            *cand = list;
            *n_cand = n;
            return RHZ_OK;
        }
        cap = n;
    }
    return set_err(RHZ_ERR_HIP, "candidate compaction did not converge");
}

// Features of the candidates (and their certificate flags).
// Features 5 and 7 (the x and y axis parity counts): the triangles rotated
// so that x, then y, is the ray axis, exact row lists of the rotated lattice
// (each triangle in the rows whose ray point its box contains), and one
// crossing count per candidate.  Lists that would exceed csr_max_entries give
// every candidate 0.5 (no answer) instead.
int axis_parity_features(Ctx& c, const int64_t* list, int64_t n, float* f) {
    const rhz_sdf_params* p = c.p;
    const int slots[2] = {5, 7};
    if (n <= 0) return RHZ_OK;
    float* rt = (float*)ws_get(c.w, S_AX_TRIS, sizeof(float) * 12 * (size_t)std::max<int64_t>(c.nt, 1));
    int4* cover = (int4*)ws_get(c.w, S_AX_COVER, sizeof(int4) * (size_t)std::max<int64_t>(c.nt, 1));
    int32_t* big = (int32_t*)ws_get(c.w, S_AX_BIG, sizeof(int32_t) * (size_t)std::max<int64_t>(c.nt, 1));
    if (!rt || !cover || !big) return set_err(RHZ_ERR_NOMEM, "axis row-list allocation failed");
    for (int axis = 0; axis < 2; ++axis) {
        const Grid r = rotated_grid(c.g, axis);
        const int64_t rows = (int64_t)r.dims[0] * r.dims[1];
        uint32_t* counts = (uint32_t*)ws_get(c.w, S_AX_COUNTS, sizeof(uint32_t) * (size_t)rows);
        int64_t* start = (int64_t*)ws_get(c.w, S_AX_START, sizeof(int64_t) * (size_t)(rows + 1));
        uint32_t* cursor = (uint32_t*)ws_get(c.w, S_AX_CURSOR, sizeof(uint32_t) * (size_t)rows);
        void* scan = ws_get(c.w, S_AX_SCAN, scan_scratch_bytes(rows));
        if (!counts || !start || !cursor || !scan) return set_err(RHZ_ERR_NOMEM, "axis row-list allocation failed");
        HIP_TRY(launch_rotate_tris(c.tris, c.nt, axis, rt, c.s));
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_AXCSR), 0, sizeof(unsigned long long), c.s));
        HIP_TRY(launch_axis_rows_count(rt, c.nt, r, p->jitter_x, p->jitter_y, cover, counts, big,
                                       cnt_ptr(c.w, C_AXBIG), cnt_ptr(c.w, C_AXCSR), c.s));
        const int rc = read_counters(c);
        if (rc) return rc;
        const int64_t entries = (int64_t)hc(c, C_AXCSR);
        if (entries > p->csr_max_entries) {
            HIP_TRY(launch_fill_feature(f, n, slots[axis], 0.5f, c.s));
            continue;
        }
        int32_t* ids = (int32_t*)ws_get(c.w, S_AX_IDS, sizeof(int32_t) * (size_t)std::max<int64_t>(entries, 1));
        if (!ids) return set_err(RHZ_ERR_NOMEM, "axis row-list allocation failed");
        HIP_TRY(launch_scan_u32(counts, rows, start, scan, c.s));
        HIP_TRY(launch_axis_rows_fill(c.nt, r, cover, big, cnt_ptr(c.w, C_AXBIG), start, cursor, ids, c.s));
        HIP_TRY(launch_axis_parity(rt, start, ids, list, n, c.g, axis, p->jitter_x, p->jitter_y, f, slots[axis], c.s));
    }
    return RHZ_OK;
}

int learned_features(Ctx& c, uint8_t* sign, const CsrBuffers& b, const int64_t* list, int64_t n, float** feats,
                     uint8_t** risk, int64_t* n_risky) {
    float* f = (float*)ws_get(c.w, S_FEAT, sizeof(float) * 12 * (size_t)std::max<int64_t>(n, 1));
    uint8_t* r = (uint8_t*)ws_get(c.w, S_RISK, (size_t)std::max<int64_t>(n, 1));
    float2* zr = (float2*)ws_get(c.w, S_ZRANGE, sizeof(float2) * (size_t)std::max<int64_t>(c.nt, 1));
    const int64_t entries = c.st->csr_entries;
    const int64_t nsup = num_supercells(c.g, c.p->sup);
    unsigned long long* keys = (unsigned long long*)ws_get(c.w, S_ZKEY, sizeof(unsigned long long) * (size_t)entries);
    unsigned long long* keys_tmp =
        (unsigned long long*)ws_get(c.w, S_ZKEY_TMP, sizeof(unsigned long long) * (size_t)entries);
    unsigned* max_ext = (unsigned*)ws_get(c.w, S_ZEXT, sizeof(unsigned));
    const size_t sort_bytes = zsort_scratch_bytes(entries, nsup, b.start);
    void* sort_scr = sort_bytes ? ws_get(c.w, S_ZSORT, sort_bytes) : nullptr;
    if (!f || !r || !zr || !keys || !keys_tmp || !max_ext || !sort_scr)
        return set_err(RHZ_ERR_NOMEM, "feature buffer allocation failed");
    HIP_TRY(launch_zsort_lists(c.tris, c.nt, b, entries, nsup, zr, max_ext, keys_tmp, keys, sort_scr, sort_bytes,
                               c.s));
    HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_RISKY), 0, sizeof(unsigned long long), c.s));
    HIP_TRY(launch_features(c.tris, c.nt, b, zr, keys, max_ext, sign, list, n, c.g, c.p->sup, c.p->jitter_x,
                            c.p->jitter_y, f, r, cnt_ptr(c.w, C_RISKY), c.s));
    int rc = axis_parity_features(c, list, n, f);
    if (rc) return rc;
    rc = read_counters(c);
    if (rc) return rc;
    *feats = f;
    *risk = r;
    *n_risky = (int64_t)hc(c, C_RISKY);
    return RHZ_OK;
}

// Ranks the candidates and returns the top ones (at least `k`, whole score
// bins) in the S_SEL list.
int learned_select(Ctx& c, const int64_t* cand, const float* feats, int64_t n, int64_t k, int64_t** sel,
                   int64_t* n_sel) {
    const rhz_sdf_params* p = c.p;
    int64_t* out = (int64_t*)ws_get(c.w, S_SEL, sizeof(int64_t) * (size_t)std::max<int64_t>(n, 1));
    if (!out) return set_err(RHZ_ERR_NOMEM, "selection buffer allocation failed");
    *sel = out;
    if (n == 0 || (k < 0 && p->scorer == 2)) {
        // nothing ranked (a blind scorer has no positives)
        *n_sel = 0;
        return RHZ_OK;
    }
    if (k >= n) {
        HIP_TRY(hipMemcpyAsync(out, cand, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToDevice, c.s));
        *n_sel = n;
        return RHZ_OK;
    }
    float* score = (float*)ws_get(c.w, S_SCORE, sizeof(float) * (size_t)n);
    unsigned* hist = (unsigned*)ws_get(c.w, S_HIST, sizeof(unsigned) * (size_t)score_bins());
    if (!score || !hist) return set_err(RHZ_ERR_NOMEM, "score buffer allocation failed");
    float lo, hi;
    const float* blob = p->mlp_weights;
#ifdef RHZ_HAVE_BUILTIN_PROPOSER
    if (!blob) blob = kProposerBlob;
#endif
    if (p->scorer == 0 && blob) {
        float* dblob = (float*)ws_get(c.w, S_MLP, sizeof(float) * (size_t)mlp_blob_count());
        if (!dblob) return set_err(RHZ_ERR_NOMEM, "proposer parameter allocation failed");
        HIP_TRY(hipMemcpyAsync(dblob, blob, sizeof(float) * (size_t)mlp_blob_count(), hipMemcpyHostToDevice, c.s));
        HIP_TRY(launch_mlp_forward(dblob, feats, n, score, c.s));
        lo = -40.0f;
        hi = 40.0f;
    } else if (p->scorer == 2 || p->scorer == 3) {
        // blind hashed subset (scores in [0, 0.5)), or the certificate flag
        // alone (flagged in [1, 1.5), the others below)
        HIP_TRY(launch_hash_score(p->scorer == 3 ? feats : nullptr, cand, n, score, c.s));
        lo = 0.0f;
        hi = p->scorer == 3 ? 1.5f : 0.5f;
    } else {
        // margin = s_out * t - d / h lies in [-0.17, 0.063] for candidates
        HIP_TRY(launch_margin_score(feats, n, score, c.s));
        lo = -0.5f;
        hi = 0.25f;
    }
    int min_bin = 0;
    if (k < 0) {
        // classifier mode: every candidate scored positive (bin of score 0;
        // for the flag ranking, the flagged ones)
        const double pos = (p->scorer == 3) ? 1.0 : 0.0;
        min_bin = (int)std::floor((pos - lo) * score_bins() / (hi - lo));
    } else {
        HIP_TRY(launch_score_hist(score, n, lo, hi, hist, c.s));
        std::vector<unsigned> hh((size_t)score_bins());
        HIP_TRY(hipMemcpyAsync(hh.data(), hist, sizeof(unsigned) * hh.size(), hipMemcpyDeviceToHost, c.s));
        HIP_TRY(hipStreamSynchronize(c.s));
        int64_t acc = 0;
        for (int bi = score_bins() - 1; bi >= 0; --bi) {
            acc += hh[(size_t)bi];
            if (acc >= k) {
                min_bin = bi;
                break;
            }
        }
    }
    HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_SEL), 0, sizeof(unsigned long long), c.s));
    HIP_TRY(launch_select(score, cand, n, lo, hi, min_bin, out, n, cnt_ptr(c.w, C_SEL), c.s));
    int rc = read_counters(c);
    if (rc) return rc;
    *n_sel = (int64_t)hc(c, C_SEL);
    return RHZ_OK;
}

// Fast path with the learned sign gate: signed scanline signs, exact
// re-probes of the top-ranked candidates, growth around every mismatch,
// then the reference far-field policy.
// The learned sign stage after the candidate list: features, ranking, the
// exact probes of the ranked and the certificate-flagged candidates and their
// commits into `sign`.  Reads the triangles, the CSR, `sign` and the
// candidates, never the distance field, so it may overlap the eikonal rounds.
int learned_sign_stage(Ctx& c, uint8_t* sign, const CsrBuffers& b, unsigned* bits, const int64_t* cand,
                       int64_t n_cand, Timer& tm, int mark_sel, int mark_end) {
    const rhz_sdf_params* p = c.p;
    rhz_sdf_stats* st = c.st;
    const int64_t V = c.g.n();
    float* feats = nullptr;
    uint8_t* risk = nullptr;
    int64_t n_risky = 0;
    int rc = learned_features(c, sign, b, cand, n_cand, &feats, &risk, &n_risky);
    if (rc) return rc;
    st->candidates = n_cand;
    st->risky_candidates = n_risky;
    const int64_t k = (p->reprobe_fraction < 0.0f)
                          ? -1
                          : std::max<int64_t>(1, (int64_t)((double)p->reprobe_fraction * (double)V));
    int64_t* sel = nullptr;
    int64_t n_sel = 0;
    rc = learned_select(c, cand, feats, n_cand, k, &sel, &n_sel);
    if (rc) return rc;
    tm.mark(mark_sel, c.s);
    HIP_TRY(hipMemsetAsync(bits, 0, sizeof(unsigned) * (size_t)((V + 31) / 32), c.s));
    HIP_TRY(launch_mark_probed(sel, n_sel, bits, c.s));
    int64_t mism_cap = std::max<int64_t>(n_sel, 1 << 16);
    int64_t* mism = (int64_t*)ws_get(c.w, S_MISM, sizeof(int64_t) * (size_t)mism_cap);
    if (!mism) return set_err(RHZ_ERR_NOMEM, "mismatch list allocation failed");
    int64_t* cur = sel;
    int64_t n_cur = n_sel;
    int gens = 0;
    st->ranked_probes = n_sel;
    const bool merged = p->learned_certify && n_risky > 0 && p->learned_max_patch_gens == 0;
    if (merged) {
        // Without growth the fallback (every flagged candidate the ranking left
        // out) does not depend on the ranked probes' outcome: both lists go
        // through one spatially sorted winding pass, committed as two ranges
        // so that their mismatches are counted apart.
        int64_t* all = (int64_t*)ws_get(c.w, S_PROBES, sizeof(int64_t) * (size_t)(n_sel + n_risky));
        if (!all) return set_err(RHZ_ERR_NOMEM, "probe list allocation failed");
        if (n_sel > 0)
            HIP_TRY(hipMemcpyAsync(all, sel, sizeof(int64_t) * (size_t)n_sel, hipMemcpyDeviceToDevice, c.s));
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_GROW), 0, sizeof(unsigned long long), c.s));
        HIP_TRY(launch_unprobed_risky(cand, risk, n_cand, bits, all + n_sel, n_risky, cnt_ptr(c.w, C_GROW), c.s));
        rc = read_counters(c);
        if (rc) return rc;
        const int64_t n_fb = std::min<int64_t>((int64_t)hc(c, C_GROW), n_risky);
        const int64_t n_all = n_sel + n_fb;
        mism = (int64_t*)ws_get(c.w, S_MISM, sizeof(int64_t) * (size_t)std::max<int64_t>(n_all, 1));
        if (!mism) return set_err(RHZ_ERR_NOMEM, "mismatch list allocation failed");
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_MISM), 0, sizeof(unsigned long long), c.s));
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_MISM2), 0, sizeof(unsigned long long), c.s));
        rc = probe_commit_ranges(c, all, n_all, n_sel, sign, mism, n_sel, mism + n_sel, n_fb);
        if (rc) return rc;
        rc = read_counters(c);
        if (rc) return rc;
        st->ranked_mismatches = (int64_t)hc(c, C_MISM);
        st->fallback_probes = n_fb;
        st->fallback_mismatches = (int64_t)hc(c, C_MISM2);
        st->mismatches_found = st->ranked_mismatches + st->fallback_mismatches;
        st->reprobed_voxels = n_all;
        n_cur = 0;
    }
    while (n_cur > 0) {
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_MISM), 0, sizeof(unsigned long long), c.s));
        rc = probe_commit(c, cur, n_cur, sign, mism, mism_cap);
        if (rc) return rc;
        st->reprobed_voxels += n_cur;
        rc = read_counters(c);
        if (rc) return rc;
        const int64_t m = (int64_t)hc(c, C_MISM);
        if (m > mism_cap) return set_err(RHZ_ERR_HIP, "mismatch list overflow");
        st->mismatches_found += m;
        if (gens == 0) st->ranked_mismatches = m;
        if (m == 0 || gens >= p->learned_max_patch_gens) break;
        // grow into the 26-neighbourhood of this generation's mismatches
        const int64_t gcap = 26 * m;
        int64_t* grown = (int64_t*)ws_get(c.w, S_SEL, sizeof(int64_t) * (size_t)std::max(gcap, n_sel));
        if (!grown) return set_err(RHZ_ERR_NOMEM, "growth list allocation failed");
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_GROW), 0, sizeof(unsigned long long), c.s));
        HIP_TRY(launch_grow(mism, m, c.g.dims, bits, grown, gcap, cnt_ptr(c.w, C_GROW), c.s));
        rc = read_counters(c);
        if (rc) return rc;
        n_cur = std::min<int64_t>((int64_t)hc(c, C_GROW), gcap);
        cur = grown;
        if (n_cur > mism_cap) {
            mism_cap = n_cur;
            mism = (int64_t*)ws_get(c.w, S_MISM, sizeof(int64_t) * (size_t)mism_cap);
            if (!mism) return set_err(RHZ_ERR_NOMEM, "mismatch list allocation failed");
        }
        ++gens;
    }
    st->patch_rounds = gens;
    if (p->learned_certify && n_risky > 0 && !merged) {
        // certificate fallback: every flagged candidate not probed yet
        int64_t* fb = (int64_t*)ws_get(c.w, S_SEL, sizeof(int64_t) * (size_t)n_risky);
        if (!fb) return set_err(RHZ_ERR_NOMEM, "fallback list allocation failed");
        HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_GROW), 0, sizeof(unsigned long long), c.s));
        HIP_TRY(launch_unprobed_risky(cand, risk, n_cand, bits, fb, n_risky, cnt_ptr(c.w, C_GROW), c.s));
        rc = read_counters(c);
        if (rc) return rc;
        const int64_t n_fb = std::min<int64_t>((int64_t)hc(c, C_GROW), n_risky);
        st->fallback_probes = n_fb;
        if (n_fb > 0) {
            if (n_fb > mism_cap) {
                mism_cap = n_fb;
                mism = (int64_t*)ws_get(c.w, S_MISM, sizeof(int64_t) * (size_t)mism_cap);
                if (!mism) return set_err(RHZ_ERR_NOMEM, "mismatch list allocation failed");
            }
            HIP_TRY(hipMemsetAsync(cnt_ptr(c.w, C_MISM), 0, sizeof(unsigned long long), c.s));
            rc = probe_commit(c, fb, n_fb, sign, mism, mism_cap);
            if (rc) return rc;
            rc = read_counters(c);
            if (rc) return rc;
            st->fallback_mismatches = (int64_t)hc(c, C_MISM);
            st->mismatches_found += st->fallback_mismatches;
            st->reprobed_voxels += n_fb;
        }
    }
    st->sites = (int32_t)std::min<int64_t>(st->mismatches_found, INT32_MAX);
    st->probes = 1;
    tm.mark(mark_end, c.s);
    return RHZ_OK;
}

// Lattices from this size on run the sign stage on the side stream while the
// eikonal rounds run (tuning key "verify_overlap").
constexpr int64_t kOverlapMinVoxels = 1ll << 24;
std::atomic<int> g_verify_overlap{1};
// 1: rhz_sdf_bake runs the eikonal rounds of all its lattices as batched
// launches (tuning key "bake_batched"); 0: every item's whole build on a lane.
std::atomic<int> g_bake_batched{1};

// Parity signs over the candidate-list CSR, then the exact fixup of the
// deferred columns.
int learned_scan(Ctx& c, uint8_t* sign, CsrBuffers* b) {
    int unknown = 0;
    int rc = run_parity(c, sign, &unknown, true, b);
    if (rc) return rc;
    c.st->unknown_flag = unknown;
    if (unknown) {
        int64_t fixed = 0;
        rc = fixup_unknown(c, sign, &fixed);
        if (rc) return rc;
        c.st->fixup_voxels = fixed;
    }
    return RHZ_OK;
}

// The two halves of a FAST_LEARNED build that the batched bake runs apart
// from the eikonal rounds.  learned_band: fill and band splat into the field
// (all the rounds need).  learned_signs: candidate compaction from the band,
// signed scan, fixup and the sign stage, into `sign` (all the clamp and the
// apply need; nothing here writes the field).
int learned_band(Ctx& c) {
    if (!int_range_ok(c.g, c.nt) || c.nt <= 0) {
        c.st->decline_reason = 1;
        return RHZ_DECLINED;
    }
    HIP_TRY(launch_fill(c.d, c.g.n(), kFar, c.s));
    HIP_TRY(launch_splat(c.tris, c.nt, c.g, (float)c.p->band_voxels * c.g.h, c.d, c.s));
    return RHZ_OK;
}

int learned_signs(Ctx& c, uint8_t* sign) {
    const int64_t V = c.g.n();
    unsigned* bits = (unsigned*)ws_get(c.w, S_BITS, sizeof(unsigned) * (size_t)((V + 31) / 32));
    if (!bits) {
        c.st->decline_reason = 4;
        return RHZ_DECLINED;
    }
    int rc = zero_counters(c);
    if (rc) return rc;
    Timer tm(true);
    tm.mark(0, c.s);
    int64_t* cand = nullptr;
    int64_t n_cand = 0;
    rc = learned_candidate_list(c, &cand, &n_cand);
    if (rc) return rc;
    tm.mark(1, c.s);
    CsrBuffers b;
    rc = learned_scan(c, sign, &b);
    if (rc) return rc;
    tm.mark(2, c.s);
    rc = learned_sign_stage(c, sign, b, bits, cand, n_cand, tm, 3, 4);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c.s));
    c.st->ms_parity = tm.ms(1, 2);
    c.st->ms_proposer = tm.ms(0, 1) + tm.ms(2, 3);
    c.st->ms_verify = tm.ms(3, 4);
    return RHZ_OK;
}

int run_learned(Ctx& c) {
    const rhz_sdf_params* p = c.p;
    rhz_sdf_stats* st = c.st;
    const int64_t V = c.g.n();
    if (!int_range_ok(c.g, c.nt) || c.nt <= 0) {
        st->decline_reason = 1;
        return RHZ_DECLINED;
    }
    uint8_t* sign = (uint8_t*)ws_get(c.w, S_SIGN, (size_t)V);
    unsigned* bits = (unsigned*)ws_get(c.w, S_BITS, sizeof(unsigned) * (size_t)((V + 31) / 32));
    if (!sign || !bits) {
        st->decline_reason = 4;
        return RHZ_DECLINED;
    }
    Timer tm(true);
    tm.mark(0, c.s);
    int rc = zero_counters(c);
    if (rc) return rc;
    HIP_TRY(launch_fill(c.d, V, kFar, c.s));
    HIP_TRY(launch_splat(c.tris, c.nt, c.g, (float)p->band_voxels * c.g.h, c.d, c.s));
    tm.mark(1, c.s);
    int64_t* cand = nullptr;
    int64_t n_cand = 0;
    CsrBuffers b;
    Workspace* sw = (g_verify_overlap.load(std::memory_order_relaxed) && V >= kOverlapMinVoxels) ? side_of(c.w) : nullptr;
    if (!sw) {
        int unknown = 0;
        rc = run_parity(c, sign, &unknown, true, &b);
        if (rc) return rc;
        tm.mark(2, c.s);
        st->unknown_flag = unknown;
        if (unknown) {
            int64_t fixed = 0;
            rc = fixup_unknown(c, sign, &fixed);
            if (rc) return rc;
            st->fixup_voxels = fixed;
        }
        tm.mark(3, c.s);
        rc = learned_candidate_list(c, &cand, &n_cand);
        if (rc) return rc;
        rc = learned_sign_stage(c, sign, b, bits, cand, n_cand, tm, 4, 5);
        if (rc) return rc;
        FsmField fld{};
        int64_t far = 0;
        rc = run_fsm_rounds(c, &far, &fld);
        if (rc) return rc;
        rc = run_fsm_finish(c, sign, far, fld, [&] { tm.mark(6, c.s); });
        if (rc) return rc;
        tm.mark(7, c.s);
        HIP_TRY(hipStreamSynchronize(c.s));
        st->ms_splat = tm.ms(0, 1);
        st->ms_parity = tm.ms(1, 2);
        st->ms_fixup = tm.ms(2, 3);
        st->ms_proposer = tm.ms(3, 4);
        st->ms_verify = tm.ms(4, 5);
        st->ms_eikonal = tm.ms(5, 6);
        st->ms_apply = tm.ms(6, 7);
        st->ms_total = tm.ms(0, 7);
        return RHZ_OK;
    }
    // Overlapped: the candidate compaction is the last step that reads the
    // band, so the main stream runs it and then the eikonal rounds; from the
    // fork on the side stream runs the scan, the fixup and the sign stage,
    // which write the signs but never touch the field; both join before the
    // clamp and the apply, which read the signs.
    rc = learned_candidate_list(c, &cand, &n_cand);
    if (rc) return rc;
    tm.mark(2, c.s);
    Ctx cv = c;
    cv.w = sw;
    cv.s = c.w->side_stream;
    cv.sweeps_t = nullptr;
    cv.trans_t = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    if (hipEventCreateWithFlags(&join, hipEventDisableTiming) != hipSuccess) {
        (void)hipEventDestroy(fork);
        return set_err(RHZ_ERR_HIP, "event creation failed");
    }
    Timer ts(true);
    int rc_side = RHZ_OK;
    std::string side_err;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipEventRecord(fork, c.s);
    if (e == hipSuccess) e = hipStreamWaitEvent(cv.s, fork, 0);
    if (e == hipSuccess) e = hipMemsetAsync(sw->d_cnt, 0, 512, cv.s);
    if (e != hipSuccess) {
        (void)hipEventDestroy(fork);
        (void)hipEventDestroy(join);
        return set_err(RHZ_ERR_HIP, "sign-stage fork failed: %s", hipGetErrorString(e));
    }
    auto side_work = [&]() {
        if (hipSetDevice(dev) != hipSuccess) {
            rc_side = set_err(RHZ_ERR_HIP, "sign stage: hipSetDevice failed");
        } else {
            ts.mark(0, cv.s);
            rc_side = learned_scan(cv, sign, &b);
            ts.mark(1, cv.s);
            if (rc_side == RHZ_OK) rc_side = learned_sign_stage(cv, sign, b, bits, cand, n_cand, ts, 2, 3);
        }
        if (rc_side) side_err = rhz_last_error();
        if (hipEventRecord(join, cv.s) != hipSuccess && rc_side == RHZ_OK) {
            rc_side = set_err(RHZ_ERR_HIP, "sign stage: event record failed");
            side_err = rhz_last_error();
        }
    };
    std::thread side;
    try {
        side = std::thread(side_work);
    } catch (...) {
        side_work();  // no thread: the stage runs first, still on its own stream
    }
    int64_t far = 0;
    FsmField fld{};
    rc = run_fsm_rounds(c, &far, &fld);
    if (side.joinable()) side.join();
    tm.mark(3, c.s);
    e = hipStreamWaitEvent(c.s, join, 0);
    (void)hipEventDestroy(fork);
    (void)hipEventDestroy(join);
    if (rc_side < 0) return set_err(rc_side, "%s", side_err.c_str());
    if (rc_side > 0) return rc_side;  // declined (decline_reason says why); the caller falls back
    if (rc) return rc;
    if (e != hipSuccess) return set_err(RHZ_ERR_HIP, "sign-stage join failed: %s", hipGetErrorString(e));
    rc = run_fsm_finish(c, sign, far, fld, [&] { tm.mark(4, c.s); });
    if (rc) return rc;
    tm.mark(5, c.s);
    HIP_TRY(hipStreamSynchronize(c.s));
    st->ms_splat = tm.ms(0, 1);
    st->ms_parity = ts.ms(0, 1);                   // scan + fixup (side)
    st->ms_fixup = 0.0;                            // (inside ms_parity here)
    st->ms_proposer = tm.ms(1, 2) + ts.ms(1, 2);  // candidate list (main) + features and ranking (side)
    st->ms_verify = ts.ms(2, 3);                   // exact probes and commits (side)
    st->ms_eikonal = tm.ms(2, 3);
    st->ms_apply = tm.ms(4, 5);
    st->ms_total = tm.ms(0, 5);
    return RHZ_OK;
}

int check_params(const rhz_sdf_params* p) {
    if (!p || p->struct_size != (int32_t)sizeof(rhz_sdf_params))
        return set_err(RHZ_ERR_ARG, "rhz_sdf_params.struct_size mismatch (rebuild against rhizome_sdf.h)");
    if (p->parity_max_hits < 1 || p->parity_max_hits > 128)
        return set_err(RHZ_ERR_ARG, "parity_max_hits must be in [1, 128]");
    if (p->sup < 1 || p->band_voxels < 0 || p->verify_samples < 1 || p->verify_gens < 1 || p->verify_gens > 3)
        return set_err(RHZ_ERR_ARG, "invalid policy constants");
    if (p->fsm_policy < RHZ_FSM_REFERENCE || p->fsm_policy > RHZ_FSM_OCTANT)
        return set_err(RHZ_ERR_ARG, "fsm_policy must be RHZ_FSM_REFERENCE, RHZ_FSM_CONVERGE or RHZ_FSM_OCTANT");
    return RHZ_OK;
}

int dispatch_inner(Ctx& c);

int dispatch(Ctx& c) {
    EventAcc sw, wd, tr, ff;
    c.sweeps_t = &sw;
    c.wind_t = &wd;
    c.trans_t = &tr;
    c.far_t = &ff;
    int rc = dispatch_inner(c);
    if (rc >= 0 && hipStreamSynchronize(c.s) == hipSuccess) {
        c.st->ms_sweeps = sw.total_ms();
        c.st->ms_winding = wd.total_ms();
        c.st->ms_transposes = tr.total_ms();
        c.st->ms_far_field = ff.total_ms();
        c.st->sweep_calls = (int64_t)(sw.used / 2);
        c.st->winding_calls = (int64_t)(wd.used / 2);
    }
    c.sweeps_t = nullptr;
    c.wind_t = nullptr;
    c.trans_t = nullptr;
    c.far_t = nullptr;
    return rc;
}

int dispatch_inner(Ctx& c) {
    const rhz_sdf_params* p = c.p;
    const int64_t V = c.g.n();
    int method = p->method;
    if (method == RHZ_METHOD_AUTO) {
        if (V >= p->fast_min_voxels) {
            int rc = run_fast(c);
            if (rc == RHZ_OK) {
                c.st->method_used = RHZ_METHOD_FAST;
                return RHZ_OK;
            }
            if (rc < 0) return rc;
            // declined: fall through to the exact kernel (:1189-1195)
        }
        c.st->method_used = RHZ_METHOD_LEGACY;
        return run_legacy(c);
    }
    c.st->method_used = method;
    switch (method) {
        case RHZ_METHOD_FAST:
            return run_fast(c);
        case RHZ_METHOD_LEGACY:
            return run_legacy(c);
        case RHZ_METHOD_PLAIN_EXACT:
            return run_plain_exact(c);
        case RHZ_METHOD_FAST_LEARNED:
            return run_learned(c);
        default:
            return set_err(RHZ_ERR_ARG, "method %d is not available", method);
    }
}

int make_grid(const int32_t min_c[3], const int32_t max_c[3], float h, Grid* g) {
    if (!min_c || !max_c) return set_err(RHZ_ERR_ARG, "null lattice bounds");
    if (!(h > 0.0f) || !std::isfinite(h)) return set_err(RHZ_ERR_ARG, "voxel size must be positive");
    for (int j = 0; j < 3; ++j) {
        int64_t d = (int64_t)max_c[j] - (int64_t)min_c[j] + 1;
        if (d <= 0 || d > 2000000000LL) return set_err(RHZ_ERR_ARG, "empty or oversized lattice on axis %d", j);
        g->mn[j] = min_c[j];
        g->dims[j] = (int32_t)d;
    }
    g->h = h;
    if (g->n() > 2000000000LL) return set_err(RHZ_ERR_ARG, "lattice exceeds 2e9 voxels");
    return RHZ_OK;
}

int need_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return set_err(RHZ_ERR_NODEV, "no HIP device available");
    }
    return RHZ_OK;
}

}  // namespace
}  // namespace rhz

using namespace rhz;

extern "C" {

void rhz_sdf_params_default(rhz_sdf_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->struct_size = (int32_t)sizeof(rhz_sdf_params);
    p->method = RHZ_METHOD_AUTO;
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
    p->fsm_policy = RHZ_FSM_REFERENCE;
    p->csr_max_entries = 100000000LL;
    p->jitter_x = 0.037f;
    p->jitter_y = 0.051f;
    p->clamp_intended_sign = 1;
    p->fsm_round_cap = 512;
    p->winding_precision = RHZ_WINDING_FP32_KAHAN;
    p->fp64_band = 1e-3f;
    p->reprobe_fraction = -1.0f;
    p->learned_max_patch_gens = 0;
    p->candidate_radius = 0.0f;
    p->scorer = 0;
    p->learned_certify = 1;
    p->far_field_beta = 4.0f;
    p->mlp_weights = nullptr;
}

const char* rhz_last_error(void) { return g_err.c_str(); }

int rhz_sdf_log_line(const rhz_sdf_stats* st, int32_t band_voxels, char* buf, int64_t cap) {
    if (!st || !buf || cap < 1) return set_err(RHZ_ERR_ARG, "null stats or buffer");
    const double total = st->ms_host_total;
    const double to_parity = st->ms_splat + st->ms_parity;
    const int n = std::snprintf(
        buf, (size_t)cap,
        "Fast SDF path OK: %.0f ms total (splat+parity %.0f ms | eikonal %d rounds %.0f ms | verify+sign+copy %.0f "
        "ms) | %dx%dx%d voxels, %lld tris, band %d vox, fixup %lld | verify %d probes, %d sites, %lld voxels "
        "patched in %d rounds | far residual %lld, extra rounds %d, safe-clamped %lld",
        total, to_parity, st->eikonal_rounds, st->ms_eikonal, total - to_parity - st->ms_eikonal, st->dims[0],
        st->dims[1], st->dims[2], (long long)st->n_triangles, band_voxels, (long long)st->fixup_voxels, st->probes,
        st->sites, (long long)st->patched_voxels, st->patch_rounds, (long long)st->far_residual, st->extra_rounds,
        (long long)st->safe_clamped);
    return n < 0 ? set_err(RHZ_ERR_ARG, "formatting failed") : RHZ_OK;
}

int rhz_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

// "<base>@<class>": a sweep tuning key of one lattice-size class (0..3).
// Returns the class, or -1 when the key has no class suffix, or -2 when the
// suffix is malformed; *base_len receives the length of the base name.
static int class_suffix(const char* key, size_t* base_len) {
    const char* at = std::strchr(key, '@');
    *base_len = at ? (size_t)(at - key) : std::strlen(key);
    if (!at) return -1;
    if (at[1] < '0' || at[1] >= '0' + kShapeClasses || at[2] != '\0') return -2;
    return at[1] - '0';
}

static bool base_is(const char* key, size_t len, const char* name) {
    return std::strlen(name) == len && !std::strncmp(key, name, len);
}

static int class_tuning(const char* key, int cls, size_t len, int64_t* get, int64_t set) {
    static const char* kShapeKeys[3] = {"sweep_shape_x", "sweep_shape_y", "sweep_shape_z"};
    static const char* kBatchKeys[3] = {"batch_shape_x", "batch_shape_y", "batch_shape_z"};
    const SweepTuning t = sweep_tuning();
    if (cls < 0) return set_err(RHZ_ERR_ARG, "malformed class suffix in tuning key '%s'", key);
    for (int b = 0; b < 2; ++b) {
        if (base_is(key, len, b ? "batch_step" : "sweep_step")) {
            if (get) {
                *get = b ? t.bcls_step[cls] : t.cls_step[cls];
                return RHZ_OK;
            }
            if (!(b ? set_batch_class_step(cls, (int)set) : set_class_step(cls, (int)set)))
                return set_err(RHZ_ERR_ARG, "%s must be 0, 2, 4 or 8", key);
            return RHZ_OK;
        }
        for (int a = 0; a < 3; ++a)
            if (base_is(key, len, b ? kBatchKeys[a] : kShapeKeys[a])) {
                if (get) {
                    *get = b ? t.bcls_shape[cls][a] : t.cls_shape[cls][a];
                    return RHZ_OK;
                }
                if (!(b ? set_batch_class_shape(cls, a, (int)set) : set_class_shape(cls, a, (int)set)))
                    return set_err(RHZ_ERR_ARG, "%s must be 0, 1, 21, 22, 12 or 14", key);
                return RHZ_OK;
            }
    }
    return set_err(RHZ_ERR_ARG, "tuning key '%s' has no per-class form", key);
}

int rhz_sweep_shape_class(const int32_t dims[3]) {
    if (!dims || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return set_err(RHZ_ERR_ARG, "dims must be positive");
    return shape_class((int64_t)dims[0] * dims[1] * dims[2]);
}

int rhz_tuning_set(const char* key, int64_t value) {
    if (!key) return set_err(RHZ_ERR_ARG, "null tuning key");
    size_t len = 0;
    const int cls = class_suffix(key, &len);
    if (cls != -1) return class_tuning(key, cls, len, nullptr, value);
    const SweepTuning t = sweep_tuning();
    if (!std::strcmp(key, "sweep_step")) {
        if (!set_sweep_tuning((int)value, t.rows, t.fmajor)) return set_err(RHZ_ERR_ARG, "sweep_step must be 2, 4 or 8");
        return RHZ_OK;
    }
    if (!std::strcmp(key, "sweep_rows")) {
        if (!set_sweep_tuning(t.step, (int)value, t.fmajor)) return set_err(RHZ_ERR_ARG, "sweep_rows must be 4, 8 or 16");
        return RHZ_OK;
    }
    if (!std::strcmp(key, "sweep_fmajor")) {
        if (!set_sweep_tuning(t.step, t.rows, (int)value)) return set_err(RHZ_ERR_ARG, "sweep_fmajor must be 0 or 1");
        return RHZ_OK;
    }
    if (!std::strcmp(key, "fsm_batch_rounds_max_voxels")) {
        if (value < 0) return set_err(RHZ_ERR_ARG, "fsm_batch_rounds_max_voxels must be >= 0");
        set_fsm_batching(value, t.fsm_graph);
        return RHZ_OK;
    }
    if (!std::strcmp(key, "fsm_graph")) {
        if (value != 0 && value != 1) return set_err(RHZ_ERR_ARG, "fsm_graph must be 0 or 1");
        set_fsm_batching(t.batch_rounds_max_voxels, (int)value);
        return RHZ_OK;
    }
    if (!std::strcmp(key, "fsm_layout")) {
        if (!set_fsm_layout((int)value)) return set_err(RHZ_ERR_ARG, "fsm_layout must be 0 or 1");
        return RHZ_OK;
    }
    if (!std::strcmp(key, "sweep_run_chunk")) {
        if (!set_run_chunk((int)value)) return set_err(RHZ_ERR_ARG, "sweep_run_chunk must be 16 or 32");
        return RHZ_OK;
    }
    if (!std::strcmp(key, "sweep_run_block")) {
        if (!set_run_block((int)value)) return set_err(RHZ_ERR_ARG, "sweep_run_block must be 0, 41 or 42");
        return RHZ_OK;
    }
    if (!std::strcmp(key, "verify_overlap")) {
        if (value != 0 && value != 1) return set_err(RHZ_ERR_ARG, "verify_overlap must be 0 or 1");
        g_verify_overlap.store((int)value, std::memory_order_relaxed);
        return RHZ_OK;
    }
    if (!std::strcmp(key, "bake_batched")) {
        if (value != 0 && value != 1) return set_err(RHZ_ERR_ARG, "bake_batched must be 0 or 1");
        g_bake_batched.store((int)value, std::memory_order_relaxed);
        return RHZ_OK;
    }
    static const char* kShapeKeys[3] = {"sweep_shape_x", "sweep_shape_y", "sweep_shape_z"};
    for (int a = 0; a < 3; ++a)
        if (!std::strcmp(key, kShapeKeys[a])) {
            if (!set_sweep_shape(a, (int)value))
                return set_err(RHZ_ERR_ARG, "%s must be 1, 21, 22, 12 or 14", kShapeKeys[a]);
            return RHZ_OK;
        }
    return set_err(RHZ_ERR_ARG, "unknown tuning key '%s'", key);
}

int rhz_tuning_get(const char* key, int64_t* value) {
    if (!key || !value) return set_err(RHZ_ERR_ARG, "null argument");
    size_t len = 0;
    const int cls = class_suffix(key, &len);
    if (cls != -1) return class_tuning(key, cls, len, value, 0);
    const SweepTuning t = sweep_tuning();
    if (!std::strcmp(key, "sweep_step")) {
        *value = t.step;
        return RHZ_OK;
    }
    if (!std::strcmp(key, "sweep_rows")) {
        *value = t.rows;
        return RHZ_OK;
    }
    if (!std::strcmp(key, "sweep_fmajor")) {
        *value = t.fmajor;
        return RHZ_OK;
    }
    static const char* kShapeKeys[3] = {"sweep_shape_x", "sweep_shape_y", "sweep_shape_z"};
    for (int a = 0; a < 3; ++a)
        if (!std::strcmp(key, kShapeKeys[a])) {
            *value = t.shape[a];
            return RHZ_OK;
        }
    if (!std::strcmp(key, "fsm_batch_rounds_max_voxels")) {
        *value = t.batch_rounds_max_voxels;
        return RHZ_OK;
    }
    if (!std::strcmp(key, "fsm_graph")) {
        *value = t.fsm_graph;
        return RHZ_OK;
    }
    if (!std::strcmp(key, "fsm_layout")) {
        *value = t.fsm_layout;
        return RHZ_OK;
    }
    if (!std::strcmp(key, "sweep_run_chunk")) {
        *value = t.run_chunk;
        return RHZ_OK;
    }
    if (!std::strcmp(key, "sweep_run_block")) {
        *value = t.run_block;
        return RHZ_OK;
    }
    if (!std::strcmp(key, "verify_overlap")) {
        *value = g_verify_overlap.load(std::memory_order_relaxed);
        return RHZ_OK;
    }
    if (!std::strcmp(key, "bake_batched")) {
        *value = g_bake_batched.load(std::memory_order_relaxed);
        return RHZ_OK;
    }
    return set_err(RHZ_ERR_ARG, "unknown tuning key '%s'", key);
}

int rhz_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
    int rc = need_device();
    if (rc) return rc;
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (uint64_t)f;
    if (total_bytes) *total_bytes = (uint64_t)t;
    return RHZ_OK;
}

// Frees every slot of w and of its side workspace, and what they retired.
static void ws_free_slots(Workspace* w) {
    ws_reap(w);
    for (Workspace* x : {w, w->side}) {
        if (!x) continue;
        for (int i = 0; i < S_COUNT; ++i) {
            if (x->p[i]) (void)hipFree(x->p[i]);
            x->p[i] = nullptr;
            x->bytes[i] = 0;
        }
    }
}

void rhz_release_workspace(void) {
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        Workspace* w = workspace();
        if (w) ws_free_slots(w);
    }
    // the bake lanes' workspaces and the batched rounds' buffers (their
    // counters and streams stay)
    std::lock_guard<std::mutex> lk(g_lane_mu);
    const int d = current_device();
    if (d < 0 || d >= 64) return;
    for (int l = 0; l < kMaxLanes; ++l)
        if (g_lanes[d][l].ready) ws_free_slots(&g_lanes[d][l].ws);
    if (g_bake[d].ready) ws_free_slots(&g_bake[d].ws);
}

// The body of rhz_sdf_compute_device; the caller holds g_ws_mu.
static int compute_device_locked(const rhz_triangle* d_tris, int64_t n_tris, const Grid& g,
                                 const rhz_sdf_params* params, float* d_out, hipStream_t stream,
                                 rhz_sdf_stats* st, std::chrono::steady_clock::time_point t0) {
    std::memset(st, 0, sizeof(*st));
    for (int j = 0; j < 3; ++j) st->dims[j] = g.dims[j];
    st->n_triangles = n_tris;
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    if (n_tris == 0) return set_err(RHZ_ERR_ARG, "no triangles");
    Ctx c{reinterpret_cast<const float*>(d_tris), n_tris, g, params, d_out, stream, st, w};
    int rc = dispatch(c);
    st->status = rc;
    st->ms_host_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int rhz_sdf_compute_device(const rhz_triangle* d_tris, int64_t n_tris, float h, const int32_t min_c[3],
                           const int32_t max_c[3], const rhz_sdf_params* params, float* d_out, void* stream,
                           rhz_sdf_stats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    int rc = need_device();
    if (rc) return rc;
    rc = check_params(params);
    if (rc) return rc;
    Grid g;
    rc = make_grid(min_c, max_c, h, &g);
    if (rc) return rc;
    if (n_tris < 0 || (n_tris > 0 && !d_tris) || !d_out) return set_err(RHZ_ERR_ARG, "null buffer");
    rhz_sdf_stats local;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    return compute_device_locked(d_tris, n_tris, g, params, d_out, (hipStream_t)stream, stats ? stats : &local, t0);
}

// ---- batched bake: the eikonal rounds of every lattice in shared launches ----
namespace rhz {
namespace {

enum BakeSlot { B_SIGN = 0, B_FSM, B_FLAGS, B_FAR, B_JOBS, B_STARTS };
constexpr int kBakeRoundStride = 32;  // flag slots per item (>= max_rounds + 1)

BakeWork* bake_work() {
    const int d = current_device();
    if (d < 0 || d >= 64) return nullptr;
    BakeWork& b = g_bake[d];
    if (!b.ready) {
        if (hipMalloc(&b.ws.d_cnt, 512) != hipSuccess) return nullptr;
        if (hipHostMalloc((void**)&b.ws.h_cnt, 512, hipHostMallocDefault) != hipSuccess) return nullptr;
        if (hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking) != hipSuccess) return nullptr;
        b.ws.init = true;
        b.ready = true;
    }
    ws_reap(&b.ws);
    return &b;
}

struct BakeItemState {
    Grid g;
    FieldLayout lx, ly;
    int64_t cells = 0;        // floats per working copy
    float* a = nullptr;       // working copies (the result ends in a, layout lx)
    float* b = nullptr;
    uint8_t* sign = nullptr;  // final signs (dense)
    int* flags = nullptr;     // per-round change flags
    int rounds = 0, extra = 0;
    int64_t far = 0;
};

// The batched rounds apply when every lattice takes the lane sweeps and the
// build is the learned method with the reference round policy.
bool bake_batchable(const rhz_bake_item* items, int n, const std::vector<Grid>& grids, const rhz_sdf_params* p) {
    if (p->method != RHZ_METHOD_FAST_LEARNED || p->fsm_policy != RHZ_FSM_REFERENCE) return false;
    if (p->max_rounds < 1 || p->max_rounds >= kBakeRoundStride) return false;
    for (int i = 0; i < n; ++i) {
        FieldLayout lx, ly;
        fsm_layouts(grids[(size_t)i], &lx, &ly);
        if (!lane_sweeps_ok(grids[(size_t)i]) || !int_range_ok(grids[(size_t)i], items[i].n_tris)) return false;
    }
    return true;
}

// One lattice's round outside the batch (audit-gated extension rounds):
// sweep_round_p's launches on the item's working copies, in place in a.
int item_round(BakeItemState& it, int* flag, hipStream_t s) {
    const Grid& g = it.g;
    HIP_TRY(launch_transpose_xy(it.a, it.b, g, true, s, it.lx, it.ly));
    HIP_TRY(launch_sweep_yx(it.b, it.a, g, +1, flag, s, it.ly));
    HIP_TRY(launch_sweep_yx(it.a, it.b, g, -1, flag, s, it.ly));
    HIP_TRY(launch_transpose_xy(it.b, it.a, g, false, s, it.ly, it.lx));
    HIP_TRY(launch_sweep(it.a, it.b, g, 1, +1, flag, s, it.lx));
    HIP_TRY(launch_sweep(it.b, it.a, g, 1, -1, flag, s, it.lx));
    HIP_TRY(launch_sweep(it.a, it.b, g, 2, +1, flag, s, it.lx));
    HIP_TRY(launch_sweep(it.b, it.a, g, 2, -1, flag, s, it.lx));
    return RHZ_OK;
}

}  // namespace
}  // namespace rhz

// The batched bake (tuning key "bake_batched", default on): phase 1 runs fill
// and splat of every item on the lanes; then the eikonal rounds of ALL
// lattices run on the bake stream as one captured graph of batched launches
// (per round: two transposes and six sweeps, each one launch over every
// lattice; a lattice whose previous round changed nothing is skipped by its
// device flag), while the lanes run each item's sign stage (candidates,
// scan, fixup, features, ranking, exact probes) into per-item sign buffers;
// last, per item, the audit, extension rounds if any, clamp and the signed
// write-back.  Fields and counters equal the single builds'.
static int bake_batched(const rhz_bake_item* items, int32_t n_items, const rhz_sdf_params* params,
                        const std::vector<Grid>& grids, const std::vector<int>& order, std::vector<Lane*>& lanes,
                        std::vector<rhz_sdf_stats>& st, std::vector<int>& rcs, std::vector<double>& lat,
                        hipEvent_t entry, int dev) {
    const auto t_start = std::chrono::steady_clock::now();
    BakeWork* bw = bake_work();
    if (!bw) return set_err(RHZ_ERR_NOMEM, "bake workspace initialisation failed");
    const hipStream_t ms = bw->stream;
    HIP_TRY(hipStreamWaitEvent(ms, entry, 0));
    const int n = n_items;
    std::vector<BakeItemState> it((size_t)n);
    size_t sign_bytes = 0, fsm_cells = 0;
    for (int i = 0; i < n; ++i) {
        BakeItemState& s = it[(size_t)i];
        s.g = grids[(size_t)i];
        fsm_layouts(s.g, &s.lx, &s.ly);
        s.cells = std::max(s.lx.plane, s.ly.plane) * (int64_t)s.g.dims[2];
        sign_bytes += (size_t)((s.g.n() + 255) / 256 * 256);
        fsm_cells += (size_t)((s.cells + 63) / 64 * 64);
    }
    uint8_t* signs = (uint8_t*)ws_get(&bw->ws, B_SIGN, sign_bytes);
    // (a tail after the last copy, as the single builds' slots have one)
    float* fsm = (float*)ws_get(&bw->ws, B_FSM, sizeof(float) * 2 * fsm_cells + (64 << 10));
    int* flags = (int*)ws_get(&bw->ws, B_FLAGS, sizeof(int) * (size_t)n * kBakeRoundStride);
    unsigned long long* far = (unsigned long long*)ws_get(&bw->ws, B_FAR, sizeof(unsigned long long) * (size_t)n);
    if (!signs || !fsm || !flags || !far) return set_err(RHZ_ERR_NOMEM, "bake buffers allocation failed");
    {
        size_t so = 0, fo = 0;
        for (int i = 0; i < n; ++i) {
            BakeItemState& s = it[(size_t)i];
            s.sign = signs + so;
            s.a = fsm + fo;
            s.b = fsm + fsm_cells + fo;
            s.flags = flags + (size_t)i * kBakeRoundStride;
            so += (size_t)((s.g.n() + 255) / 256 * 256);
            fo += (size_t)((s.cells + 63) / 64 * 64);
        }
    }
    // job tables: transposes (round 0 from the dense band, later rounds from
    // a; the way back), sweeps x+/x- (y-fastest), y+/y-, z+/z-.  Jobs are
    // ordered by lattice-size class (shape_class): a sweep launch covers the
    // jobs of one class with that class's tile shape, and its block offsets
    // (starts) begin at 0 for the class's first job.
    std::vector<int> pos;  // item of every job position
    int cb[kShapeClasses + 1] = {0};
    for (int c = 0; c < kShapeClasses; ++c) {
        cb[c] = (int)pos.size();
        for (int i = 0; i < n; ++i)
            if (shape_class(it[(size_t)i].g.n()) == c) pos.push_back(i);
    }
    cb[kShapeClasses] = n;
    const int nT = 3, nS = 6;
    std::vector<TransJob> tj((size_t)nT * n);
    std::vector<SweepJob> sj((size_t)nS * n);
    // transposes to y-fastest / back: n + 1 offsets each; sweeps per
    // orientation and class: (jobs of the class) + 1 offsets
    const size_t sweep_base = (size_t)2 * (n + 1);
    auto sw_off = [&](int o, int c) { return sweep_base + (size_t)o * (n + kShapeClasses) + (size_t)(cb[c] + c); };
    std::vector<unsigned> starts(sweep_base + (size_t)3 * (n + kShapeClasses));
    unsigned ntr[2] = {0, 0};
    unsigned nsw[3][kShapeClasses] = {};
    for (int q = 0; q < n; ++q) {
        const int i = pos[(size_t)q];
        const BakeItemState& s = it[(size_t)i];
        const int cls = shape_class(s.g.n());
        const int dx = s.g.dims[0], dy = s.g.dims[1], dz = s.g.dims[2];
        const FieldLayout dense = dense_xy(s.g);
        int tr = 0, tc = 0;
        const int64_t b0 = transpose_batch_blocks(dy, dx, dz, &tr, &tc);
        tj[(size_t)0 * n + q] = TransJob{items[i].d_out, s.b, dense.pitch, dense.plane, s.ly.pitch, s.ly.plane,
                                        dy, dx, tr, tc, s.flags};
        tj[(size_t)1 * n + q] = TransJob{s.a, s.b, s.lx.pitch, s.lx.plane, s.ly.pitch, s.ly.plane, dy, dx, tr, tc,
                                        s.flags};
        const int64_t b1 = transpose_batch_blocks(dx, dy, dz, &tr, &tc);
        tj[(size_t)2 * n + q] = TransJob{s.b, s.a, s.ly.pitch, s.ly.plane, s.lx.pitch, s.lx.plane, dx, dy, tr, tc,
                                        s.flags};
        int nbf = 0;
        const int64_t bx = sweep_batch_blocks(0, cls, dy, dz, &nbf);
        sj[(size_t)0 * n + q] = SweepJob{s.b, s.a, s.ly.pitch, s.ly.plane, dy, dx, dz, nbf, s.g.h, s.flags};
        sj[(size_t)1 * n + q] = SweepJob{s.a, s.b, s.ly.pitch, s.ly.plane, dy, dx, dz, nbf, s.g.h, s.flags};
        const int64_t by = sweep_batch_blocks(1, cls, dx, dz, &nbf);
        sj[(size_t)2 * n + q] = SweepJob{s.a, s.b, s.lx.pitch, s.lx.plane, dx, dy, dz, nbf, s.g.h, s.flags};
        sj[(size_t)3 * n + q] = SweepJob{s.b, s.a, s.lx.pitch, s.lx.plane, dx, dy, dz, nbf, s.g.h, s.flags};
        const int64_t bz = sweep_batch_blocks(2, cls, dx, dy, &nbf);
        sj[(size_t)4 * n + q] = SweepJob{s.a, s.b, s.lx.plane, s.lx.pitch, dx, dz, dy, nbf, s.g.h, s.flags};
        sj[(size_t)5 * n + q] = SweepJob{s.b, s.a, s.lx.plane, s.lx.pitch, dx, dz, dy, nbf, s.g.h, s.flags};
        const int64_t bt[2] = {b0, b1};
        for (int k = 0; k < 2; ++k) {
            starts[(size_t)k * (n + 1) + q] = ntr[k];
            if ((int64_t)ntr[k] + bt[k] >= (int64_t)1 << 31) return set_err(RHZ_ERR_ARG, "batch too large");
            ntr[k] += (unsigned)bt[k];
        }
        const int64_t bs[3] = {bx, by, bz};
        for (int o = 0; o < 3; ++o) {
            starts[sw_off(o, cls) + (size_t)(q - cb[cls])] = nsw[o][cls];
            if ((int64_t)nsw[o][cls] + bs[o] >= (int64_t)1 << 31) return set_err(RHZ_ERR_ARG, "batch too large");
            nsw[o][cls] += (unsigned)bs[o];
        }
    }
    for (int k = 0; k < 2; ++k) starts[(size_t)k * (n + 1) + n] = ntr[k];
    for (int o = 0; o < 3; ++o)
        for (int c = 0; c < kShapeClasses; ++c) starts[sw_off(o, c) + (size_t)(cb[c + 1] - cb[c])] = nsw[o][c];
    const size_t tj_bytes = sizeof(TransJob) * tj.size(), sj_bytes = sizeof(SweepJob) * sj.size();
    char* jobs = (char*)ws_get(&bw->ws, B_JOBS, tj_bytes + sj_bytes);
    unsigned* d_starts = (unsigned*)ws_get(&bw->ws, B_STARTS, sizeof(unsigned) * starts.size());
    if (!jobs || !d_starts) return set_err(RHZ_ERR_NOMEM, "bake tables allocation failed");
    HIP_TRY(hipMemcpyAsync(jobs, tj.data(), tj_bytes, hipMemcpyHostToDevice, ms));
    HIP_TRY(hipMemcpyAsync(jobs + tj_bytes, sj.data(), sj_bytes, hipMemcpyHostToDevice, ms));
    HIP_TRY(hipMemcpyAsync(d_starts, starts.data(), sizeof(unsigned) * starts.size(), hipMemcpyHostToDevice, ms));
    const TransJob* d_tj = reinterpret_cast<const TransJob*>(jobs);
    const SweepJob* d_sj = reinterpret_cast<const SweepJob*>(jobs + tj_bytes);
    auto tr_starts = [&](int k) { return d_starts + (size_t)k * (n + 1); };
    // the sweeps of orientation o, direction dir, over every class with jobs
    auto sweep_all = [&](int o, int dir, int r) -> hipError_t {
        const int tab = 2 * o + (dir > 0 ? 0 : 1);
        for (int c = 0; c < kShapeClasses; ++c) {
            const int cnt = cb[c + 1] - cb[c];
            if (cnt == 0) continue;
            const hipError_t e = launch_sweep_batch(o, c, dir, d_sj + (size_t)tab * n + cb[c], d_starts + sw_off(o, c),
                                                    cnt, nsw[o][c], r, ms);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };

    // phase 1 (lanes): fill and splat of every item
    const int n_lanes = (int)lanes.size();
    std::atomic<int> next{0};
    std::vector<std::string> errs((size_t)n);
    auto run_items = [&](int l, bool signs_phase) {
        (void)hipSetDevice(dev);
        Lane* L = lanes[(size_t)l];
        for (;;) {
            const int k = next.fetch_add(1);
            if (k >= n) break;
            const int i = order[(size_t)k];
            if (rcs[(size_t)i] != RHZ_OK) continue;
            Ctx c{reinterpret_cast<const float*>(items[i].d_tris), items[i].n_tris, grids[(size_t)i], params,
                  items[i].d_out, L->stream, &st[(size_t)i], &L->ws};
            const int r = signs_phase ? learned_signs(c, it[(size_t)i].sign) : learned_band(c);
            if (r != RHZ_OK) {
                rcs[(size_t)i] = r;
                if (r < 0) errs[(size_t)i] = rhz_last_error();
            }
        }
        if (hipStreamSynchronize(L->stream) != hipSuccess) {
            (void)hipGetLastError();
            for (int i = 0; i < n; ++i)
                if (rcs[(size_t)i] == RHZ_OK) rcs[(size_t)i] = set_err(RHZ_ERR_HIP, "bake lane sync failed");
        }
    };
    // every lane (the calling thread is lane 0) takes items until none is left
    auto run_lanes = [&](bool signs_phase) {
        next.store(0);
        std::vector<std::thread> th;
        for (int l = 1; l < n_lanes; ++l) th.emplace_back(run_items, l, signs_phase);
        run_items(0, signs_phase);
        for (auto& t : th) t.join();
    };
    run_lanes(false);
    const auto t_band = std::chrono::steady_clock::now();

    // phase 2: the rounds of every lattice as one captured graph on the bake
    // stream, while the lanes run the sign stages
    const int cap = params->max_rounds;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = RHZ_OK;
    {
        hipError_t e = hipStreamBeginCapture(ms, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) e = launch_zero_i32(flags, n * kBakeRoundStride, ms);
        for (int r = 0; r < cap && e == hipSuccess; ++r) {
            e = launch_transpose_batch(d_tj + (size_t)(r == 0 ? 0 : 1) * n, tr_starts(0), n, ntr[0], r, ms);
            if (e == hipSuccess) e = sweep_all(0, +1, r);
            if (e == hipSuccess) e = sweep_all(0, -1, r);
            if (e == hipSuccess) e = launch_transpose_batch(d_tj + (size_t)2 * n, tr_starts(1), n, ntr[1], r, ms);
            if (e == hipSuccess) e = sweep_all(1, +1, r);
            if (e == hipSuccess) e = sweep_all(1, -1, r);
            if (e == hipSuccess) e = sweep_all(2, +1, r);
            if (e == hipSuccess) e = sweep_all(2, -1, r);
        }
        const hipError_t ec = hipStreamEndCapture(ms, &graph);
        if (e == hipSuccess) e = ec;
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphLaunch(exec, ms);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            rc = set_err(RHZ_ERR_HIP, "batched rounds: %s", hipGetErrorString(e));
        }
    }
    run_lanes(true);  // the sign stages, beside the rounds
    if (rc == RHZ_OK && hipStreamSynchronize(ms) != hipSuccess) rc = set_err(RHZ_ERR_HIP, "batched rounds failed");
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    const auto t_rounds = std::chrono::steady_clock::now();

    // rounds per lattice, then the audit of every lattice
    std::vector<int> hflags((size_t)n * kBakeRoundStride);
    HIP_TRY(hipMemcpy(hflags.data(), flags, sizeof(int) * hflags.size(), hipMemcpyDeviceToHost));
    HIP_TRY(launch_zero_i32(reinterpret_cast<int*>(far), 2 * n, ms));
    for (int i = 0; i < n; ++i) {
        BakeItemState& s = it[(size_t)i];
        s.rounds = cap;
        for (int r = 0; r < cap; ++r)
            if (hflags[(size_t)i * kBakeRoundStride + r] == 0) {
                s.rounds = r + 1;
                break;
            }
        HIP_TRY(launch_count_far_rows(s.a, s.lx.pitch, s.lx.plane, s.g.dims[0], s.g.dims[1], s.g.dims[2], far + i, ms));
    }
    std::vector<unsigned long long> hfar((size_t)n);
    HIP_TRY(hipMemcpyAsync(hfar.data(), far, sizeof(unsigned long long) * (size_t)n, hipMemcpyDeviceToHost, ms));
    HIP_TRY(hipStreamSynchronize(ms));
    // audit-gated extension rounds (rare: a round reaches every voxel that
    // straight lines connect to the band)
    for (int i = 0; i < n; ++i) {
        BakeItemState& s = it[(size_t)i];
        s.far = (int64_t)hfar[(size_t)i];
        while (s.far > 0 && s.extra < params->max_extra_rounds) {
            int* ext_flag = s.flags + kBakeRoundStride - 1;
            rc = item_round(s, ext_flag, ms);
            if (rc) return rc;
            ++s.extra;
            HIP_TRY(launch_zero_i32(reinterpret_cast<int*>(far + i), 2, ms));
            HIP_TRY(launch_count_far_rows(s.a, s.lx.pitch, s.lx.plane, s.g.dims[0], s.g.dims[1], s.g.dims[2], far + i,
                                          ms));
            HIP_TRY(hipMemcpyAsync(&hfar[(size_t)i], far + i, sizeof(unsigned long long), hipMemcpyDeviceToHost, ms));
            HIP_TRY(hipStreamSynchronize(ms));
            s.far = (int64_t)hfar[(size_t)i];
        }
    }
    // clamp and the signed write-back
    for (int i = 0; i < n; ++i) {
        if (rcs[(size_t)i] != RHZ_OK) continue;
        BakeItemState& s = it[(size_t)i];
        const FieldLayout dense = dense_xy(s.g);
        const int dx = s.g.dims[0], dy = s.g.dims[1], dz = s.g.dims[2];
        rhz_sdf_stats& si = st[(size_t)i];
        si.eikonal_rounds = s.rounds + s.extra;
        si.extra_rounds = s.extra;
        si.sweeps = 6 * (s.rounds + s.extra);
        si.safe_clamped = 0;
        si.far_residual = 0;
        if (s.far == 0) {
            HIP_TRY(launch_copy_apply_rows(s.a, s.lx.pitch, s.lx.plane, items[i].d_out, s.sign, dx, dy, dz, ms));
        } else {
            HIP_TRY(launch_copy_rows(s.a, s.lx.pitch, s.lx.plane, items[i].d_out, dense.pitch, dense.plane, dx, dy,
                                     dz, ms));
            si.safe_clamped = s.far;
            HIP_TRY(launch_clamp_far(items[i].d_out, s.sign, s.g.n(), (float)params->band_voxels * s.g.h,
                                     params->clamp_intended_sign, ms));
            HIP_TRY(launch_zero_i32(reinterpret_cast<int*>(far + i), 2, ms));
            HIP_TRY(launch_count_far(items[i].d_out, s.g.n(), far + i, ms));
            HIP_TRY(hipMemcpyAsync(&hfar[(size_t)i], far + i, sizeof(unsigned long long), hipMemcpyDeviceToHost, ms));
            HIP_TRY(hipStreamSynchronize(ms));
            si.far_residual = (int64_t)hfar[(size_t)i];
            HIP_TRY(launch_apply(items[i].d_out, s.sign, s.g.n(), ms));
        }
    }
    HIP_TRY(hipStreamSynchronize(ms));
    const auto t_end = std::chrono::steady_clock::now();
    const double ms_all = std::chrono::duration<double, std::milli>(t_end - t_start).count();
    const double ms_band = std::chrono::duration<double, std::milli>(t_band - t_start).count();
    const double ms_rounds = std::chrono::duration<double, std::milli>(t_rounds - t_band).count();
    for (int i = 0; i < n; ++i) {
        if (rcs[(size_t)i] < 0 && !errs[(size_t)i].empty()) set_err(rcs[(size_t)i], "%s", errs[(size_t)i].c_str());
        rhz_sdf_stats& si = st[(size_t)i];
        si.status = rcs[(size_t)i];
        si.method_used = RHZ_METHOD_FAST_LEARNED;
        si.ms_splat = ms_band;
        si.ms_eikonal = ms_rounds;  // the shared rounds (sign stages beside them)
        si.ms_total = ms_all;
        si.ms_host_total = ms_all;
        lat[(size_t)i] = ms_all;  // every field is complete when the batch is
    }
    return RHZ_OK;
}

int rhz_sdf_bake(const rhz_bake_item* items, int32_t n_items, const rhz_sdf_params* params, int32_t n_lanes,
                 rhz_sdf_stats* stats, double* ms_item, void* stream) {
    int rc = need_device();
    if (rc) return rc;
    rc = check_params(params);
    if (rc) return rc;
    if (n_items < 0 || (n_items > 0 && !items)) return set_err(RHZ_ERR_ARG, "null item list");
    if (n_lanes < 1 || n_lanes > kMaxLanes) return set_err(RHZ_ERR_ARG, "n_lanes must be in [1, %d]", kMaxLanes);
    std::vector<Grid> grids((size_t)n_items);
    for (int i = 0; i < n_items; ++i) {
        rc = make_grid(items[i].min_c, items[i].max_c, items[i].h, &grids[(size_t)i]);
        if (rc) return rc;
        if (items[i].n_tris <= 0 || !items[i].d_tris || !items[i].d_out)
            return set_err(RHZ_ERR_ARG, "item %d: null or empty buffer", i);
    }
    // largest lattices first: the long items start early, the short ones fill in
    std::vector<int> order((size_t)n_items);
    for (int i = 0; i < n_items; ++i) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return grids[(size_t)a].n() > grids[(size_t)b].n(); });
    std::lock_guard<std::mutex> lk(g_lane_mu);
    int dev = current_device();
    std::vector<Lane*> lanes((size_t)n_lanes);
    for (int l = 0; l < n_lanes; ++l) {
        lanes[(size_t)l] = bake_lane(l);
        if (!lanes[(size_t)l]) return set_err(RHZ_ERR_NOMEM, "bake lane %d initialisation failed", l);
    }
    // Every lane starts after the work the caller queued on `stream` (its
    // inputs may still be in flight there); the call returns only when every
    // item is complete, so work queued on `stream` afterwards follows the bake.
    hipEvent_t entry = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&entry, hipEventDisableTiming));
    struct EventGuard {
        hipEvent_t e;
        ~EventGuard() { (void)hipEventDestroy(e); }
    } entry_guard{entry};
    HIP_TRY(hipEventRecord(entry, (hipStream_t)stream));
    for (Lane* L : lanes) HIP_TRY(hipStreamWaitEvent(L->stream, entry, 0));
    std::vector<rhz_sdf_stats> st((size_t)n_items);
    std::vector<double> lat((size_t)n_items, 0.0);
    std::vector<int> rcs((size_t)n_items, RHZ_OK);
    if (g_bake_batched.load() && n_items > 0 && bake_batchable(items, n_items, grids, params)) {
        for (int i = 0; i < n_items; ++i) {
            std::memset(&st[(size_t)i], 0, sizeof(rhz_sdf_stats));
            for (int j = 0; j < 3; ++j) st[(size_t)i].dims[j] = grids[(size_t)i].dims[j];
            st[(size_t)i].n_triangles = items[i].n_tris;
        }
        rc = bake_batched(items, n_items, params, grids, order, lanes, st, rcs, lat, entry, dev);
        for (Lane* L : lanes) ws_reap(&L->ws);
        if (rc) return rc;
        int first = RHZ_OK;
        for (int i = 0; i < n_items; ++i) {
            if (stats) stats[i] = st[(size_t)i];
            if (ms_item) ms_item[i] = lat[(size_t)i];
            if (rcs[(size_t)i] < 0 && first == RHZ_OK) first = rcs[(size_t)i];
        }
        if (first < 0) return set_err(first, "a bake item failed (status in its stats)");
        return RHZ_OK;
    }
    std::atomic<int> next{0};
    auto worker = [&](int l) {
        (void)hipSetDevice(dev);
        Lane* L = lanes[(size_t)l];
        for (;;) {
            const int k = next.fetch_add(1);
            if (k >= n_items) break;
            const int i = order[(size_t)k];
            const auto t0 = std::chrono::steady_clock::now();
            rhz_sdf_stats& si = st[(size_t)i];
            std::memset(&si, 0, sizeof(si));
            const Grid& g = grids[(size_t)i];
            for (int j = 0; j < 3; ++j) si.dims[j] = g.dims[j];
            si.n_triangles = items[i].n_tris;
            Ctx c{reinterpret_cast<const float*>(items[i].d_tris), items[i].n_tris, g, params, items[i].d_out,
                  L->stream, &si, &L->ws};
            int r = dispatch(c);
            if (r >= 0 && hipStreamSynchronize(L->stream) != hipSuccess) r = set_err(RHZ_ERR_HIP, "bake sync failed");
            si.status = r;
            rcs[(size_t)i] = r;
            lat[(size_t)i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            si.ms_host_total = lat[(size_t)i];
        }
    };
    std::vector<std::thread> th;
    for (int l = 1; l < n_lanes; ++l) th.emplace_back(worker, l);
    worker(0);
    for (auto& t : th) t.join();
    for (Lane* L : lanes) ws_reap(&L->ws);  // every item has synchronised its lane
    int first = RHZ_OK;
    for (int i = 0; i < n_items; ++i) {
        if (stats) stats[i] = st[(size_t)i];
        if (ms_item) ms_item[i] = lat[(size_t)i];
        if (rcs[(size_t)i] < 0 && first == RHZ_OK) first = rcs[(size_t)i];
    }
    if (first < 0) return set_err(first, "a bake item failed (status in its stats)");
    return RHZ_OK;
}

int rhz_sdf_compute(const rhz_triangle* tris, int64_t n_tris, float h, const int32_t min_c[3],
                    const int32_t max_c[3], const rhz_sdf_params* params, float* out_host, rhz_sdf_stats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    int rc = need_device();
    if (rc) return rc;
    rc = check_params(params);
    if (rc) return rc;
    Grid g;
    rc = make_grid(min_c, max_c, h, &g);
    if (rc) return rc;
    if (n_tris <= 0 || !tris || !out_host) return set_err(RHZ_ERR_ARG, "null or empty buffer");
    rhz_sdf_stats local;
    rhz_sdf_stats* st = stats ? stats : &local;
    // One lock over staging, compute and the copy back: the staged slots are
    // shared with every other host entry point of this device.
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    float* d_tris = (float*)ws_get(w, S_TRIS, sizeof(rhz_triangle) * (size_t)n_tris);
    float* d_out = (float*)ws_get(w, S_OUT, sizeof(float) * (size_t)g.n());
    if (!d_tris || !d_out) return set_err(RHZ_ERR_NOMEM, "device buffer allocation failed");
    HIP_TRY(hipMemcpy(d_tris, tris, sizeof(rhz_triangle) * (size_t)n_tris, hipMemcpyHostToDevice));
    rc = compute_device_locked(reinterpret_cast<const rhz_triangle*>(d_tris), n_tris, g, params, d_out, nullptr, st,
                               t0);
    if (rc == RHZ_OK) HIP_TRY(hipMemcpy(out_host, d_out, sizeof(float) * (size_t)g.n(), hipMemcpyDeviceToHost));
    st->ms_host_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

int rhz_build_triangles(const float* verts, int64_t n_verts, const uint32_t* idx, int64_t n_tris, rhz_triangle* out,
                        int64_t* n_out) {
    int rc = need_device();
    if (rc) return rc;
    if (!verts || !idx || !out || !n_out || n_verts < 0 || n_tris < 0) return set_err(RHZ_ERR_ARG, "null buffer");
    *n_out = 0;
    if (n_tris == 0) return RHZ_OK;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    float* dv = (float*)ws_get(w, S_MISC, sizeof(float) * 3 * (size_t)std::max<int64_t>(n_verts, 1));
    uint32_t* di = (uint32_t*)ws_get(w, S_MISC2, sizeof(uint32_t) * 3 * (size_t)n_tris);
    uint32_t* valid = (uint32_t*)ws_get(w, S_COUNTS, sizeof(uint32_t) * (size_t)n_tris);
    int64_t* pos = (int64_t*)ws_get(w, S_START, sizeof(int64_t) * (size_t)(n_tris + 1));
    void* scan = ws_get(w, S_SCAN, scan_scratch_bytes(n_tris));
    float* dt = (float*)ws_get(w, S_TRIS, sizeof(rhz_triangle) * (size_t)n_tris);
    if (!dv || !di || !valid || !pos || !scan || !dt) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    HIP_TRY(hipMemcpy(dv, verts, sizeof(float) * 3 * (size_t)n_verts, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di, idx, sizeof(uint32_t) * 3 * (size_t)n_tris, hipMemcpyHostToDevice));
    HIP_TRY(launch_tri_valid(dv, n_verts, di, n_tris, valid, nullptr));
    HIP_TRY(launch_scan_u32(valid, n_tris, pos, scan, nullptr));
    HIP_TRY(launch_tri_emit(dv, di, n_tris, valid, pos, dt, nullptr));
    int64_t k = 0;
    HIP_TRY(hipMemcpy(&k, pos + n_tris, sizeof(int64_t), hipMemcpyDeviceToHost));
    if (k > 0) HIP_TRY(hipMemcpy(out, dt, sizeof(rhz_triangle) * (size_t)k, hipMemcpyDeviceToHost));
    *n_out = k;
    return RHZ_OK;
}

int rhz_mesh_bbox(const float* verts, int64_t n_verts, float bmin[3], float bmax[3]) {
    int rc = need_device();
    if (rc) return rc;
    if (!verts || !bmin || !bmax || n_verts < 0) return set_err(RHZ_ERR_ARG, "null buffer");
    if (n_verts == 0) {
        for (int j = 0; j < 3; ++j) bmin[j] = bmax[j] = 0.0f;  // cuda_mesh_ops.cu:206-210
        return RHZ_OK;
    }
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    int nb = bbox_blocks(n_verts);
    float* dv = (float*)ws_get(w, S_MISC, sizeof(float) * 3 * (size_t)n_verts);
    float* part = (float*)ws_get(w, S_MISC2, sizeof(float) * 6 * (size_t)nb + 64);
    if (!dv || !part) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    HIP_TRY(hipMemcpy(dv, verts, sizeof(float) * 3 * (size_t)n_verts, hipMemcpyHostToDevice));
    HIP_TRY(launch_bbox(dv, n_verts, part, nb, part + 6 * nb, nullptr));
    float r[6];
    HIP_TRY(hipMemcpy(r, part + 6 * nb, sizeof(r), hipMemcpyDeviceToHost));
    for (int j = 0; j < 3; ++j) {
        bmin[j] = r[j];
        bmax[j] = r[3 + j];
    }
    return RHZ_OK;
}

int rhz_frame_grid(const float* verts, int64_t n_verts, float h, int32_t min_c[3], int32_t max_c[3]) {
    if (!(h > 0.0f)) return set_err(RHZ_ERR_ARG, "voxel size must be positive");
    float bmin[3], bmax[3];
    int rc = rhz_mesh_bbox(verts, n_verts, bmin, bmax);
    if (rc) return rc;
    float pad = h * 10.0f;
    for (int j = 0; j < 3; ++j) {
        float lo = bmin[j] - pad, hi = bmax[j] + pad;
        min_c[j] = (int32_t)std::floor(lo / h);
        max_c[j] = (int32_t)std::ceil(hi / h);
    }
    return RHZ_OK;
}

int rhz_winding_numbers(const float* verts, int64_t n_verts, const uint32_t* idx, int64_t n_tris, const float* pts,
                        int64_t n_pts, float* out_w) {
    int rc = need_device();
    if (rc) return rc;
    if (n_pts < 0 || n_tris < 0 || (n_pts > 0 && (!pts || !out_w))) return set_err(RHZ_ERR_ARG, "null buffer");
    if (n_pts == 0) return RHZ_OK;
    if (n_tris == 0) {
        for (int64_t i = 0; i < n_pts; ++i) out_w[i] = 0.0f;  // cuda_mesh_ops.cu:834-838
        return RHZ_OK;
    }
    for (int64_t t = 0; t < 3 * n_tris; ++t)
        if ((int64_t)idx[t] >= n_verts) return set_err(RHZ_ERR_ARG, "index out of range");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    float* dv = (float*)ws_get(w, S_MISC, sizeof(float) * 3 * (size_t)n_verts);
    uint32_t* di = (uint32_t*)ws_get(w, S_MISC2, sizeof(uint32_t) * 3 * (size_t)n_tris);
    float* dp = (float*)ws_get(w, S_TRIS, sizeof(float) * 3 * (size_t)n_pts);
    float* dw = (float*)ws_get(w, S_OUT, sizeof(float) * (size_t)n_pts);
    if (!dv || !di || !dp || !dw) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    HIP_TRY(hipMemcpy(dv, verts, sizeof(float) * 3 * (size_t)n_verts, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di, idx, sizeof(uint32_t) * 3 * (size_t)n_tris, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dp, pts, sizeof(float) * 3 * (size_t)n_pts, hipMemcpyHostToDevice));
    HIP_TRY(launch_winding_points(dv, di, n_tris, dp, n_pts, dw, nullptr));
    HIP_TRY(hipMemcpy(out_w, dw, sizeof(float) * (size_t)n_pts, hipMemcpyDeviceToHost));
    return RHZ_OK;
}

// Stages an indexed mesh in the workspace (vertices in S_MISC, indices in
// S_MISC2); rejects out-of-range indices.
static int stage_mesh(Workspace* w, const float* verts, int64_t n_verts, const uint32_t* idx, int64_t n_tris,
                      float** dv, uint32_t** di) {
    for (int64_t t = 0; t < 3 * n_tris; ++t)
        if ((int64_t)idx[t] >= n_verts) return set_err(RHZ_ERR_ARG, "index out of range");
    *dv = (float*)ws_get(w, S_MISC, sizeof(float) * 3 * (size_t)std::max<int64_t>(n_verts, 1));
    *di = (uint32_t*)ws_get(w, S_MISC2, sizeof(uint32_t) * 3 * (size_t)n_tris);
    if (!*dv || !*di) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    HIP_TRY(hipMemcpy(*dv, verts, sizeof(float) * 3 * (size_t)n_verts, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(*di, idx, sizeof(uint32_t) * 3 * (size_t)n_tris, hipMemcpyHostToDevice));
    return RHZ_OK;
}

// Signed volume of a staged mesh: fixed-order sum of the per-block partials.
static int staged_volume(Workspace* w, const float* dv, const uint32_t* di, int64_t n_tris, double* vol) {
    const int nb = volume_blocks(n_tris);
    double* part = (double*)ws_get(w, S_TP_VOL, sizeof(double) * (size_t)nb);
    if (!part) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    HIP_TRY(launch_volume_partials(dv, di, n_tris, part, nullptr));
    std::vector<double> h((size_t)nb);
    HIP_TRY(hipMemcpy(h.data(), part, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost));
    double sum = 0.0;
    for (double x : h) sum += x;
    *vol = sum / 6.0;
    return RHZ_OK;
}

int rhz_signed_volume(const float* verts, int64_t n_verts, const uint32_t* idx, int64_t n_tris, double* volume) {
    int rc = need_device();
    if (rc) return rc;
    if (!volume || n_tris < 0 || n_verts < 0 || (n_tris > 0 && (!idx || !verts)))
        return set_err(RHZ_ERR_ARG, "null buffer");
    *volume = 0.0;
    if (n_tris == 0 || n_verts == 0) return RHZ_OK;  // cuda_mesh_topology.cu:660-664
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    float* dv = nullptr;
    uint32_t* di = nullptr;
    rc = stage_mesh(w, verts, n_verts, idx, n_tris, &dv, &di);
    if (rc) return rc;
    return staged_volume(w, dv, di, n_tris, volume);
}

int rhz_orient_mesh(const float* verts, int64_t n_verts, uint32_t* idx, int64_t n_tris, int32_t* is_orientable,
                    int32_t* had_inconsistent, int32_t* was_inverted, double* signed_volume) {
    int rc = need_device();
    if (rc) return rc;
    if (n_tris < 0 || n_verts < 0 || (n_tris > 0 && (!idx || !verts))) return set_err(RHZ_ERR_ARG, "null buffer");
    if (n_tris >= ((int64_t)1 << 31)) return set_err(RHZ_ERR_ARG, "more than 2^31 - 1 triangles");
    int32_t orientable = 1, inconsistent = 0, inverted = 0;
    double vol = 0.0;
    if (n_tris > 0 && n_verts > 0) {  // an empty mesh is oriented (cuda_mesh_topology.cu:722-725)
        std::lock_guard<std::mutex> lk(g_ws_mu);
        Workspace* w = workspace();
        if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
        float* dv = nullptr;
        uint32_t* di = nullptr;
        rc = stage_mesh(w, verts, n_verts, idx, n_tris, &dv, &di);
        if (rc) return rc;
        const int64_t ts = edge_table_size(n_tris);
        EdgeTable E{(unsigned long long*)ws_get(w, S_TP_KEY, sizeof(unsigned long long) * (size_t)ts),
                    (uint32_t*)ws_get(w, S_TP_CNT, sizeof(uint32_t) * (size_t)ts),
                    (unsigned long long*)ws_get(w, S_TP_LO, sizeof(unsigned long long) * (size_t)ts),
                    (unsigned long long*)ws_get(w, S_TP_HI, sizeof(unsigned long long) * (size_t)ts), ts};
        int64_t* slot = (int64_t*)ws_get(w, S_TP_SLOT, sizeof(int64_t) * 3 * (size_t)n_tris);
        int32_t* nb = (int32_t*)ws_get(w, S_TP_NB, sizeof(int32_t) * 3 * (size_t)n_tris);
        uint8_t* same = (uint8_t*)ws_get(w, S_TP_SAME, 3 * (size_t)n_tris);
        int32_t* p0 = (int32_t*)ws_get(w, S_TP_P0, sizeof(int32_t) * (size_t)n_tris);
        int32_t* p1 = (int32_t*)ws_get(w, S_TP_P1, sizeof(int32_t) * (size_t)n_tris);
        uint8_t* q0 = (uint8_t*)ws_get(w, S_TP_Q0, (size_t)n_tris);
        uint8_t* q1 = (uint8_t*)ws_get(w, S_TP_Q1, (size_t)n_tris);
        if (!E.key || !E.count || !E.lo || !E.hi || !slot || !nb || !same || !p0 || !p1 || !q0 || !q1)
            return set_err(RHZ_ERR_NOMEM, "allocation failed");
        int* flags = flag_ptr(w, 0);  // [0] inconsistent, [1] changed, [2] non-orientable
        HIP_TRY(launch_zero_i32(flags, 3, nullptr));
        HIP_TRY(launch_edge_links(di, n_tris, E, slot, nb, same, flags + 0, nullptr));
        HIP_TRY(launch_label_init(n_tris, p0, q0, nullptr));
        // hook + jump steps until no parent moves (checked every 8 steps);
        // the labels settle within diameter + 1 steps
        constexpr int kBatch = 8;
        for (int64_t done = 0;; done += kBatch) {
            if (done > n_tris + kBatch) return set_err(RHZ_ERR_HIP, "component labels did not settle");
            HIP_TRY(launch_zero_i32(flags + 1, 1, nullptr));
            for (int k = 0; k < kBatch; ++k) HIP_TRY(launch_label_step(n_tris, nb, same, p0, q0, p1, q1, flags + 1, nullptr));
            int changed = 1;
            HIP_TRY(hipMemcpy(&changed, flags + 1, sizeof(int), hipMemcpyDeviceToHost));
            if (!changed) break;
        }
        HIP_TRY(launch_orient_check(n_tris, nb, same, q0, flags + 2, nullptr));
        int h[3] = {0, 0, 0};
        HIP_TRY(hipMemcpy(h, flags, sizeof(h), hipMemcpyDeviceToHost));
        inconsistent = h[0] != 0;
        orientable = h[2] == 0;
        if (inconsistent) HIP_TRY(launch_flip(di, n_tris, q0, nullptr));
        rc = staged_volume(w, dv, di, n_tris, &vol);
        if (rc) return rc;
        if (vol < -1e-5) {  // cuda_mesh_topology.cu:806-810
            HIP_TRY(launch_flip(di, n_tris, nullptr, nullptr));
            vol = -vol;
            inverted = 1;
        }
        HIP_TRY(hipMemcpy(idx, di, sizeof(uint32_t) * 3 * (size_t)n_tris, hipMemcpyDeviceToHost));
    }
    if (is_orientable) *is_orientable = orientable;
    if (had_inconsistent) *had_inconsistent = inconsistent;
    if (was_inverted) *was_inverted = inverted;
    if (signed_volume) *signed_volume = vol;
    return RHZ_OK;
}

int rhz_query_distances(const float* verts, int64_t n_verts, const uint32_t* idx, int64_t n_tris, const float* pts,
                        int64_t n_pts, float* out_d) {
    int rc = need_device();
    if (rc) return rc;
    if (n_pts < 0 || n_tris < 0 || n_verts < 0 || (n_pts > 0 && (!pts || !out_d)) ||
        (n_tris > 0 && (!verts || !idx)))
        return set_err(RHZ_ERR_ARG, "null buffer");
    if (n_pts == 0) return RHZ_OK;
    if (n_tris == 0) {
        for (int64_t i = 0; i < n_pts; ++i) out_d[i] = 0.0f;  // cuda_mesh_ops.cu:714-718
        return RHZ_OK;
    }
    for (int64_t t = 0; t < 3 * n_tris; ++t)
        if ((int64_t)idx[t] >= n_verts) return set_err(RHZ_ERR_ARG, "index out of range");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    float* dv = (float*)ws_get(w, S_MISC, sizeof(float) * 3 * (size_t)n_verts);
    uint32_t* di = (uint32_t*)ws_get(w, S_MISC2, sizeof(uint32_t) * 3 * (size_t)n_tris);
    float* dp = (float*)ws_get(w, S_TRIS, sizeof(float) * 3 * (size_t)n_pts);
    float* dd = (float*)ws_get(w, S_OUT, sizeof(float) * (size_t)n_pts);
    if (!dv || !di || !dp || !dd) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    HIP_TRY(hipMemcpy(dv, verts, sizeof(float) * 3 * (size_t)n_verts, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(di, idx, sizeof(uint32_t) * 3 * (size_t)n_tris, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dp, pts, sizeof(float) * 3 * (size_t)n_pts, hipMemcpyHostToDevice));
    HIP_TRY(launch_point_distances(dv, di, n_tris, dp, n_pts, dd, nullptr));
    HIP_TRY(hipMemcpy(out_d, dd, sizeof(float) * (size_t)n_pts, hipMemcpyDeviceToHost));
    return RHZ_OK;
}

// Marching cubes of a device field (extract_iso_surface, cuda_marching_cubes.cu:
// 829-962): the counts always, the vertices and triangles when the buffers
// hold them and both counts are positive.
int iso_surface_run(Workspace* w, const float* d_grid, const int32_t dims[3], float h, const float origin[3], float iso,
                    float* d_verts, int64_t verts_cap, uint32_t* d_tris, int64_t tris_cap, int64_t* n_verts,
                    int64_t* n_tris, hipStream_t s) {
    *n_verts = *n_tris = 0;
    if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2) return RHZ_OK;  // :842-845
    const int64_t V = (int64_t)dims[0] * dims[1] * dims[2];
    const int64_t nb = mc_blocks(dims);
    if (nb > (int64_t)INT32_MAX) return set_err(RHZ_ERR_ARG, "lattice too large");
    uint8_t* info = (uint8_t*)ws_get(w, S_MC_INFO, (size_t)V);
    uint8_t* in = (uint8_t*)ws_get(w, S_MC_IN, (size_t)V);
    uint32_t* blk = (uint32_t*)ws_get(w, S_MC_BLK, sizeof(uint32_t) * 2 * (size_t)nb);
    int64_t* start = (int64_t*)ws_get(w, S_MC_START, sizeof(int64_t) * 2 * (size_t)(nb + 1));
    void* scan = ws_get(w, S_MC_SCAN, scan_scratch_bytes(nb));
    if (!info || !in || !blk || !start || !scan)
        return set_err(RHZ_ERR_NOMEM, "marching-cubes scratch allocation failed");
    HIP_TRY(launch_mc_classify(d_grid, dims, h, iso, in, info, blk, blk + nb, s));
    HIP_TRY(launch_scan_u32(blk, nb, start, scan, s));
    HIP_TRY(launch_scan_u32(blk + nb, nb, start + nb + 1, scan, s));
    int64_t tot[2];
    HIP_TRY(hipMemcpyAsync(&tot[0], start + nb, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&tot[1], start + 2 * nb + 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *n_verts = tot[0];
    *n_tris = tot[1];
    if (tot[0] > (int64_t)UINT32_MAX) return set_err(RHZ_ERR_ARG, "more than 2^32 marching-cubes vertices");
    if (tot[0] == 0 || tot[1] == 0 || !d_verts || !d_tris || verts_cap < tot[0] || tris_cap < tot[1]) return RHZ_OK;
    uint32_t* voff = (uint32_t*)ws_get(w, S_MC_VOFF, sizeof(uint32_t) * (size_t)V);
    if (!voff) return set_err(RHZ_ERR_NOMEM, "marching-cubes scratch allocation failed");
    HIP_TRY(launch_mc_emit(d_grid, dims, h, origin, iso, in, info, start, start + nb + 1, voff, d_verts, d_tris, s));
    return RHZ_OK;
}

int rhz_iso_surface_device(const float* d_grid, const int32_t dims[3], float h, const float origin[3], float iso,
                           float* d_verts, int64_t verts_cap, uint32_t* d_tris, int64_t tris_cap, int64_t* n_verts,
                           int64_t* n_tris, void* stream) {
    int rc = need_device();
    if (rc) return rc;
    if (!d_grid || !dims || !origin || !n_verts || !n_tris) return set_err(RHZ_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    return iso_surface_run(w, d_grid, dims, h, origin, iso, d_verts, verts_cap, d_tris, tris_cap, n_verts, n_tris,
                           (hipStream_t)stream);
}

int rhz_iso_surface(const float* grid, const int32_t dims[3], float h, const float origin[3], float iso, float* verts,
                    int64_t verts_cap, uint32_t* tris, int64_t tris_cap, int64_t* n_verts, int64_t* n_tris) {
    int rc = need_device();
    if (rc) return rc;
    if (!grid || !dims || !origin || !n_verts || !n_tris) return set_err(RHZ_ERR_ARG, "null argument");
    *n_verts = *n_tris = 0;
    if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2) return RHZ_OK;
    const int64_t V = (int64_t)dims[0] * dims[1] * dims[2];
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    float* dg = (float*)ws_get(w, S_OUT, sizeof(float) * (size_t)V);
    if (!dg) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    HIP_TRY(hipMemcpy(dg, grid, sizeof(float) * (size_t)V, hipMemcpyHostToDevice));
    rc = iso_surface_run(w, dg, dims, h, origin, iso, nullptr, 0, nullptr, 0, n_verts, n_tris, nullptr);
    if (rc) return rc;
    const int64_t nv = *n_verts, nt = *n_tris;
    if (nv == 0 || nt == 0 || !verts || !tris || verts_cap < nv || tris_cap < nt) return RHZ_OK;
    float* dv = (float*)ws_get(w, S_MC_VERTS, sizeof(float) * 3 * (size_t)nv);
    uint32_t* dt = (uint32_t*)ws_get(w, S_MC_TRIS, sizeof(uint32_t) * 3 * (size_t)nt);
    if (!dv || !dt) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    rc = iso_surface_run(w, dg, dims, h, origin, iso, dv, nv, dt, nt, n_verts, n_tris, nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(verts, dv, sizeof(float) * 3 * (size_t)nv, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(tris, dt, sizeof(uint32_t) * 3 * (size_t)nt, hipMemcpyDeviceToHost));
    return RHZ_OK;
}

int interior_count_run(Workspace* w, const float* d_grid, int64_t n, int64_t* count, hipStream_t s) {
    HIP_TRY(hipMemsetAsync(cnt_ptr(w, C_INSIDE), 0, sizeof(unsigned long long), s));
    HIP_TRY(launch_interior_count(d_grid, n, cnt_ptr(w, C_INSIDE), s));
    unsigned long long c = 0;
    HIP_TRY(hipMemcpyAsync(&c, cnt_ptr(w, C_INSIDE), sizeof(c), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *count = (int64_t)c;
    return RHZ_OK;
}

int rhz_interior_count_device(const float* d_grid, int64_t n, int64_t* count, void* stream) {
    int rc = need_device();
    if (rc) return rc;
    if (!count || n < 0 || (n > 0 && !d_grid)) return set_err(RHZ_ERR_ARG, "invalid argument");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    return interior_count_run(w, d_grid, n, count, (hipStream_t)stream);
}

int rhz_selftest_sweep_sqrt(uint64_t* mismatches) {
    int rc = need_device();
    if (rc) return rc;
    if (!mismatches) return set_err(RHZ_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    unsigned long long* bad = cnt_ptr(w, C_MISM);
    HIP_TRY(hipMemsetAsync(bad, 0, sizeof(unsigned long long), nullptr));
    HIP_TRY(launch_sqrt_selftest(bad, nullptr));
    unsigned long long h = 0;
    HIP_TRY(hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost));
    *mismatches = (uint64_t)h;
    return RHZ_OK;
}

int rhz_interior_count(const float* grid, int64_t n, int64_t* count) {
    int rc = need_device();
    if (rc) return rc;
    if (!count || n < 0 || (n > 0 && !grid)) return set_err(RHZ_ERR_ARG, "invalid argument");
    *count = 0;
    if (n == 0) return RHZ_OK;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    float* dg = (float*)ws_get(w, S_OUT, sizeof(float) * (size_t)n);
    if (!dg) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    HIP_TRY(hipMemcpy(dg, grid, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    return interior_count_run(w, dg, n, count, nullptr);
}

int rhz_eikonal_redblack(float* grid, const int32_t dims[3], float h, int32_t iters) {
    int rc = need_device();
    if (rc) return rc;
    if (!grid || !dims || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0 || iters < 0)
        return set_err(RHZ_ERR_ARG, "invalid grid");
    int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    float* dg = (float*)ws_get(w, S_OUT, sizeof(float) * (size_t)n);
    if (!dg) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    HIP_TRY(hipMemcpy(dg, grid, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    for (int it = 0; it < iters; ++it) {
        HIP_TRY(launch_redblack(dg, dims, h, 0, nullptr));
        HIP_TRY(launch_redblack(dg, dims, h, 1, nullptr));
    }
    HIP_TRY(hipMemcpy(grid, dg, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    return RHZ_OK;
}

// ---- stage-level entry points -------------------------------------------------
int rhz_stage_fill(float* d, int64_t n, float value, void* stream) {
    HIP_TRY(launch_fill(d, n, value, (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_stage_splat(const rhz_triangle* d_tris, int64_t n_tris, float h, float band_width, const int32_t min_c[3],
                    const int32_t dims[3], float* d, void* stream) {
    Grid g;
    for (int j = 0; j < 3; ++j) {
        g.mn[j] = min_c[j];
        g.dims[j] = dims[j];
    }
    g.h = h;
    HIP_TRY(launch_splat(reinterpret_cast<const float*>(d_tris), n_tris, g, band_width, d, (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_stage_parity(const rhz_triangle* d_tris, int64_t n_tris, float h, float jitter_x, float jitter_y,
                     int32_t max_hits, int32_t sup, int64_t csr_max_entries, const int32_t min_c[3],
                     const int32_t dims[3], uint8_t* d_sign, int32_t* unknown_flag, int64_t* csr_entries,
                     void* stream) {
    rhz_sdf_params p;
    rhz_sdf_params_default(&p);
    p.jitter_x = jitter_x;
    p.jitter_y = jitter_y;
    p.parity_max_hits = max_hits;
    p.sup = sup;
    p.csr_max_entries = csr_max_entries;
    if (max_hits < 1 || max_hits > 128 || sup < 1) return set_err(RHZ_ERR_ARG, "invalid parity constants");
    rhz_sdf_stats st;
    std::memset(&st, 0, sizeof(st));
    Grid g;
    for (int j = 0; j < 3; ++j) {
        g.mn[j] = min_c[j];
        g.dims[j] = dims[j];
    }
    g.h = h;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    Ctx c{reinterpret_cast<const float*>(d_tris), n_tris, g, &p, nullptr, (hipStream_t)stream, &st, w};
    int rc = zero_counters(c);
    if (rc) return rc;
    int flag = 0;
    rc = run_parity(c, d_sign, &flag);
    if (unknown_flag) *unknown_flag = flag;
    if (csr_entries) *csr_entries = st.csr_entries;
    return rc;
}

int rhz_stage_fixup(const rhz_triangle* d_tris, int64_t n_tris, float h, const int32_t min_c[3],
                    const int32_t dims[3], int32_t precision, uint8_t* d_sign, int64_t* fixed, void* stream) {
    rhz_sdf_params p;
    rhz_sdf_params_default(&p);
    p.winding_precision = precision;
    rhz_sdf_stats st;
    std::memset(&st, 0, sizeof(st));
    Grid g;
    for (int j = 0; j < 3; ++j) {
        g.mn[j] = min_c[j];
        g.dims[j] = dims[j];
    }
    g.h = h;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    Ctx c{reinterpret_cast<const float*>(d_tris), n_tris, g, &p, nullptr, (hipStream_t)stream, &st, w};
    int64_t k = 0;
    int rc = fixup_unknown(c, d_sign, &k);
    if (fixed) *fixed = k;
    if (rc == RHZ_OK) HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return rc;
}

int rhz_stage_sweep(const float* src, float* dst, const int32_t dims[3], float h, int32_t axis, int32_t dir,
                    int32_t* changed, void* stream) {
    if (axis < 0 || axis > 2 || (dir != 1 && dir != -1)) return set_err(RHZ_ERR_ARG, "invalid sweep");
    Grid g;
    for (int j = 0; j < 3; ++j) {
        g.mn[j] = 0;
        g.dims[j] = dims[j];
    }
    g.h = h;
    HIP_TRY(launch_sweep(src, dst, g, axis, dir, changed, (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_stage_sweep_octant(float* d, const int32_t dims[3], float h, int32_t sx, int32_t sy, int32_t sz,
                           int32_t* changed, void* stream) {
    if ((sx != 1 && sx != -1) || (sy != 1 && sy != -1) || (sz != 1 && sz != -1))
        return set_err(RHZ_ERR_ARG, "orderings are +1 / -1 per axis");
    if (!dims || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return set_err(RHZ_ERR_ARG, "dims must be positive");
    Grid g;
    for (int j = 0; j < 3; ++j) {
        g.mn[j] = 0;
        g.dims[j] = dims[j];
    }
    g.h = h;
    HIP_TRY(launch_sweep_octant(d, g, sx, sy, sz, changed, (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_stage_sweep_round(float* a, float* b, const int32_t dims[3], float h, int32_t* changed, void* stream) {
    rhz_sdf_params p;
    rhz_sdf_params_default(&p);
    rhz_sdf_stats st;
    std::memset(&st, 0, sizeof(st));
    Grid g;
    for (int j = 0; j < 3; ++j) {
        g.mn[j] = 0;
        g.dims[j] = dims[j];
    }
    g.h = h;
    Ctx c{nullptr, 0, g, &p, a, (hipStream_t)stream, &st, nullptr};
    return sweep_round(c, a, b, changed);
}

int rhz_stage_count_far(const float* d, int64_t n, int64_t* count, void* stream) {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemsetAsync(w->d_cnt + C_FAR, 0, sizeof(unsigned long long), s));
    HIP_TRY(launch_count_far(d, n, w->d_cnt + C_FAR, s));
    unsigned long long v = 0;
    HIP_TRY(hipMemcpyAsync(&v, w->d_cnt + C_FAR, sizeof(v), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (count) *count = (int64_t)v;
    return RHZ_OK;
}

int rhz_stage_clamp_far(float* d, const uint8_t* sign, int64_t n, float band_width, int32_t intended_sign,
                        void* stream) {
    HIP_TRY(launch_clamp_far(d, sign, n, band_width, intended_sign, (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_stage_verify(const rhz_triangle* d_tris, int64_t n_tris, float h, const int32_t min_c[3],
                     const int32_t dims[3], int32_t samples, int32_t gen, int32_t precision, const uint8_t* d_sign,
                     int64_t* d_sites, int32_t* m, void* stream) {
    if (samples < 1 || gen < 0 || gen > 2) return set_err(RHZ_ERR_ARG, "invalid probe");
    rhz_sdf_params p;
    rhz_sdf_params_default(&p);
    p.winding_precision = precision;
    rhz_sdf_stats st;
    std::memset(&st, 0, sizeof(st));
    Grid g;
    for (int j = 0; j < 3; ++j) {
        g.mn[j] = min_c[j];
        g.dims[j] = dims[j];
    }
    g.h = h;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    Ctx c{reinterpret_cast<const float*>(d_tris), n_tris, g, &p, nullptr, (hipStream_t)stream, &st, w};
    int64_t* idx = (int64_t*)ws_get(w, S_SAMPLES, sizeof(int64_t) * (size_t)samples);
    double* wv = (double*)ws_get(w, S_MISC, sizeof(double) * (size_t)samples);
    void* scr = ws_get(w, S_WSCR, winding_scratch_bytes(samples, n_tris));
    int64_t* f64 = (int64_t*)ws_get(w, S_FP64, sizeof(int64_t) * (size_t)samples);
    if (!idx || !wv || !scr || !f64) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    int rc = zero_counters(c);
    if (rc) return rc;
    HIP_TRY(launch_verify_indices(g.n(), samples, gen, idx, c.s));
    HIP_TRY(launch_winding(c.tris, n_tris, g, idx, 0, samples, precision, p.fp64_band, wv, scr, f64,
                           cnt_ptr(w, C_FP64), c.s));
    if (precision != 0) {
        rc = read_counters(c);
        if (rc) return rc;
        HIP_TRY(launch_winding_fp64(c.tris, n_tris, g, idx, 0, f64, (int64_t)hc(c, C_FP64), wv, c.s));
    }
    HIP_TRY(launch_compare_samples(wv, idx, samples, d_sign, d_sites, samples, cnt_ptr(w, C_MISM), c.s));
    rc = read_counters(c);
    if (rc) return rc;
    if (m) *m = (int32_t)hc(c, C_MISM);
    return RHZ_OK;
}

int rhz_stage_mark_patch(uint8_t* d_sign, const int64_t* d_sites, int32_t n_sites, int32_t radius,
                         const int32_t dims[3], void* stream) {
    HIP_TRY(launch_mark_patch(d_sign, d_sites, n_sites, radius, dims, (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_stage_apply(float* d, const uint8_t* sign, int64_t n, void* stream) {
    HIP_TRY(launch_apply(d, sign, n, (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_stage_winding_voxels(const rhz_triangle* d_tris, int64_t n_tris, float h, const int32_t min_c[3],
                             const int32_t dims[3], const int64_t* d_voxels, int64_t n_voxels, int32_t precision,
                             double* d_w, void* stream) {
    rhz_sdf_params p;
    rhz_sdf_params_default(&p);
    p.winding_precision = precision;
    rhz_sdf_stats st;
    std::memset(&st, 0, sizeof(st));
    Grid g;
    for (int j = 0; j < 3; ++j) {
        g.mn[j] = min_c[j];
        g.dims[j] = dims[j];
    }
    g.h = h;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    Ctx c{reinterpret_cast<const float*>(d_tris), n_tris, g, &p, nullptr, (hipStream_t)stream, &st, w};
    void* scr = ws_get(w, S_WSCR, winding_scratch_bytes(n_voxels, n_tris));
    int64_t* f64 = (int64_t*)ws_get(w, S_FP64, sizeof(int64_t) * (size_t)std::max<int64_t>(n_voxels, 1));
    if (!scr || !f64) return set_err(RHZ_ERR_NOMEM, "allocation failed");
    int rc = zero_counters(c);
    if (rc) return rc;
    HIP_TRY(launch_winding(c.tris, n_tris, g, d_voxels, 0, n_voxels, precision, p.fp64_band, d_w, scr, f64,
                           cnt_ptr(w, C_FP64), c.s));
    if (precision != 0) {
        rc = read_counters(c);
        if (rc) return rc;
        HIP_TRY(launch_winding_fp64(c.tris, n_tris, g, d_voxels, 0, f64, (int64_t)hc(c, C_FP64), d_w, c.s));
    }
    HIP_TRY(hipStreamSynchronize(c.s));
    return RHZ_OK;
}

// Jittered stratified voxel ids: sample i lies in [i V / n, (i + 1) V / n),
// at a position drawn from a splitmix64 hash of (seed, i).
static void accuracy_sample_ids(int64_t V, int64_t n, uint64_t seed, std::vector<int64_t>& ids) {
    ids.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        int64_t v = (int64_t)(((double)i + u) * (double)V / (double)n);
        ids[(size_t)i] = std::min<int64_t>(std::max<int64_t>(v, 0), V - 1);
    }
}

int rhz_sdf_accuracy(const rhz_triangle* d_tris, int64_t n_tris, float h, const int32_t min_c[3],
                     const int32_t max_c[3], const float* d_field, int64_t n_samples, int32_t band_voxels,
                     uint64_t seed, void* stream, rhz_accuracy* out) {
    int rc = need_device();
    if (rc) return rc;
    Grid g;
    rc = make_grid(min_c, max_c, h, &g);
    if (rc) return rc;
    if (!out || !d_field || n_tris <= 0 || !d_tris) return set_err(RHZ_ERR_ARG, "null or empty buffer");
    if (n_samples < 1 || n_samples > (int64_t)1 << 26) return set_err(RHZ_ERR_ARG, "n_samples must be in [1, 2^26]");
    if (band_voxels < 0) return set_err(RHZ_ERR_ARG, "band_voxels must be >= 0");
    const int64_t n = std::min<int64_t>(n_samples, g.n());
    std::vector<int64_t> ids;
    accuracy_sample_ids(g.n(), n, seed, ids);
    rhz_sdf_params p;
    rhz_sdf_params_default(&p);
    p.winding_precision = RHZ_WINDING_FP32_KAHAN;
    rhz_sdf_stats st;
    std::memset(&st, 0, sizeof(st));
    std::vector<float> f((size_t)n);
    std::vector<double> d2((size_t)n), wv((size_t)n);
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        Workspace* w = workspace();
        if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
        Ctx c{reinterpret_cast<const float*>(d_tris), n_tris, g, &p, nullptr, (hipStream_t)stream, &st, w};
        int64_t* d_ids = (int64_t*)ws_get(w, S_SAMPLES, sizeof(int64_t) * (size_t)n);
        double* d_w = (double*)ws_get(w, S_MISC, sizeof(double) * (size_t)n);
        double* d_d2 = (double*)ws_get(w, S_MISC2, sizeof(double) * (size_t)n);
        float* d_f = (float*)ws_get(w, S_ACC, sizeof(float) * (size_t)n);
        void* scr = ws_get(w, S_WSCR, winding_scratch_bytes(n, n_tris));
        int64_t* f64 = (int64_t*)ws_get(w, S_FP64, sizeof(int64_t) * (size_t)n);
        if (!d_ids || !d_w || !d_d2 || !d_f || !scr || !f64) return set_err(RHZ_ERR_NOMEM, "allocation failed");
        HIP_TRY(hipMemcpyAsync(d_ids, ids.data(), sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice, c.s));
        HIP_TRY(launch_gather_f32(d_field, d_ids, n, d_f, c.s));
        HIP_TRY(launch_distance_f64(c.tris, n_tris, g, d_ids, n, d_d2, c.s));
        rc = zero_counters(c);
        if (rc) return rc;
        HIP_TRY(launch_winding(c.tris, n_tris, g, d_ids, 0, n, p.winding_precision, p.fp64_band, d_w, scr, f64,
                               cnt_ptr(w, C_FP64), c.s));
        rc = read_counters(c);
        if (rc) return rc;
        HIP_TRY(launch_winding_fp64(c.tris, n_tris, g, d_ids, 0, f64, (int64_t)hc(c, C_FP64), d_w, c.s));
        HIP_TRY(hipMemcpyAsync(f.data(), d_f, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c.s));
        HIP_TRY(hipMemcpyAsync(d2.data(), d_d2, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c.s));
        HIP_TRY(hipMemcpyAsync(wv.data(), d_w, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c.s));
        HIP_TRY(hipStreamSynchronize(c.s));
    }
    std::memset(out, 0, sizeof(*out));
    out->samples = n;
    const double hd = (double)h, band = (double)band_voxels * (double)h;
    double bs = 0, bq = 0, fs = 0, fq = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double truth = std::sqrt(d2[(size_t)i]);
        const float v = f[(size_t)i];
        if (truth <= 1e-5 * hd) {
            ++out->on_surface;  // w is 1/2 to rounding there: the sign is not checked
        } else if (std::signbit(v) != (std::fabs(wv[(size_t)i]) >= 0.5)) {
            ++out->sign_mismatches;
        }
        const double mag = std::fabs((double)v);
        if (!(mag < 1e5)) {  // the unreached marker (or NaN) is not a distance
            ++out->unreached;
            continue;
        }
        const double e = (mag - truth) / hd;
        if (truth <= band) {
            ++out->band_samples;
            bs += e;
            bq += e * e;
            out->band_max_err = std::max(out->band_max_err, std::fabs(e));
        } else {
            ++out->far_samples;
            fs += e;
            fq += e * e;
            out->far_max_err = std::max(out->far_max_err, std::fabs(e));
        }
    }
    if (out->band_samples) {
        out->band_mean_err = bs / (double)out->band_samples;
        out->band_rms_err = std::sqrt(bq / (double)out->band_samples);
    }
    if (out->far_samples) {
        out->far_mean_err = fs / (double)out->far_samples;
        out->far_rms_err = std::sqrt(fq / (double)out->far_samples);
    }
    return RHZ_OK;
}

int rhz_mlp_param_count(void) { return mlp_param_count(); }
int rhz_mlp_blob_count(void) { return mlp_blob_count(); }

int rhz_mlp_builtin(float* out) {
#ifdef RHZ_HAVE_BUILTIN_PROPOSER
    if (out) std::memcpy(out, kProposerBlob, sizeof(float) * (size_t)mlp_blob_count());
    return RHZ_OK;
#else
    (void)out;
    return RHZ_DECLINED;
#endif
}

int rhz_mlp_forward(const float* d_blob, const float* d_feats, int64_t n, float* d_logits, void* stream) {
    if (n < 0 || (n > 0 && (!d_blob || !d_feats || !d_logits))) return set_err(RHZ_ERR_ARG, "null buffer");
    HIP_TRY(launch_mlp_forward(d_blob, d_feats, n, d_logits, (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_mlp_grad(const float* d_blob, const float* d_feats, const float* d_labels, int64_t n, float pos_weight,
                 float* d_grads, float* d_loss, void* stream) {
    if (n < 0 || (n > 0 && (!d_blob || !d_feats || !d_labels || !d_grads || !d_loss)))
        return set_err(RHZ_ERR_ARG, "null buffer");
    HIP_TRY(launch_mlp_grad(d_blob, d_feats, d_labels, n, pos_weight, d_grads, d_loss, (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_mlp_adam(float* d_blob, const float* d_grads, float* d_m, float* d_v, int32_t n_params, float lr,
                 float beta1, float beta2, float eps, int32_t step, float grad_scale, float weight_decay,
                 void* stream) {
    if (n_params <= 0 || n_params > mlp_param_count() || step < 1) return set_err(RHZ_ERR_ARG, "invalid Adam step");
    HIP_TRY(launch_adam(d_blob, d_grads, d_m, d_v, n_params, lr, beta1, beta2, eps, step, grad_scale, weight_decay,
                        (hipStream_t)stream));
    return RHZ_OK;
}

int rhz_learned_candidates(const rhz_triangle* d_tris, int64_t n_tris, float h, const int32_t min_c[3],
                           const int32_t max_c[3], const rhz_sdf_params* params, int64_t* d_voxels, float* d_feats,
                           uint8_t* d_risk, uint8_t* d_scan, int64_t cap, int64_t* n_out, void* stream) {
    int rc = need_device();
    if (rc) return rc;
    rc = check_params(params);
    if (rc) return rc;
    Grid g;
    rc = make_grid(min_c, max_c, h, &g);
    if (rc) return rc;
    if (n_tris <= 0 || !d_tris || !n_out) return set_err(RHZ_ERR_ARG, "null buffer");
    if (!int_range_ok(g, n_tris)) return set_err(RHZ_ERR_ARG, "lattice out of range");
    rhz_sdf_stats st;
    std::memset(&st, 0, sizeof(st));
    std::lock_guard<std::mutex> lk(g_ws_mu);
    Workspace* w = workspace();
    if (!w) return set_err(RHZ_ERR_NOMEM, "workspace initialisation failed");
    const int64_t V = g.n();
    float* d = (float*)ws_get(w, S_OUT, sizeof(float) * (size_t)V);
    uint8_t* sign = (uint8_t*)ws_get(w, S_SIGN, (size_t)V);
    if (!d || !sign) return set_err(RHZ_ERR_NOMEM, "grid allocation failed");
    Ctx c{reinterpret_cast<const float*>(d_tris), n_tris, g, params, d, (hipStream_t)stream, &st, w};
    rc = zero_counters(c);
    if (rc) return rc;
    HIP_TRY(launch_fill(d, V, kFar, c.s));
    HIP_TRY(launch_splat(c.tris, n_tris, g, (float)params->band_voxels * h, d, c.s));
    int unknown = 0;
    CsrBuffers b;
    rc = run_parity(c, sign, &unknown, true, &b);
    if (rc) return rc;
    if (unknown) {
        int64_t fixed = 0;
        rc = fixup_unknown(c, sign, &fixed);
        if (rc) return rc;
    }
    int64_t* cand = nullptr;
    float* feats = nullptr;
    uint8_t* risk = nullptr;
    int64_t n = 0, n_risky = 0;
    rc = learned_candidate_list(c, &cand, &n);
    if (rc) return rc;
    rc = learned_features(c, sign, b, cand, n, &feats, &risk, &n_risky);
    if (rc) return rc;
    const int64_t k = std::min(n, cap);
    if (k > 0 && d_voxels) HIP_TRY(hipMemcpyAsync(d_voxels, cand, sizeof(int64_t) * (size_t)k, hipMemcpyDeviceToDevice, c.s));
    if (k > 0 && d_feats) HIP_TRY(hipMemcpyAsync(d_feats, feats, sizeof(float) * 12 * (size_t)k, hipMemcpyDeviceToDevice, c.s));
    if (k > 0 && d_risk) HIP_TRY(hipMemcpyAsync(d_risk, risk, (size_t)k, hipMemcpyDeviceToDevice, c.s));
    if (d_scan) HIP_TRY(hipMemcpyAsync(d_scan, sign, (size_t)V, hipMemcpyDeviceToDevice, c.s));
    HIP_TRY(hipStreamSynchronize(c.s));
    *n_out = n;
    return RHZ_OK;
}

}  // extern "C"
