// k_topology.hip -- winding unification and orientation of an indexed mesh
// (the pre-pass the reference runs before building an SDF,
// engine/io/obj_parser.cpp:8-34 -> CudaRasterizer::orient_mesh_gpu,
// engine/mesh/cuda_mesh_topology.cu:253-454, 707-842).
//
// Semantics kept: an undirected edge shared by exactly two triangles links
// them; it is "inconsistent" when both traverse it in the same direction.
// Every connected component is unified to the winding of its lowest-numbered
// triangle; a component whose links cannot all be satisfied is
// non-orientable; afterwards the whole mesh is inverted if its signed volume
// is below -1e-5.
//
// Built deterministically:
//  * edge table: open addressing on the 64-bit key (min vertex, max vertex);
//    per key only order-free aggregates are kept (a count, and the minimum
//    and maximum of (triangle << 1 | direction) over the triangles on the
//    edge), so the pair on a two-triangle edge does not depend on which
//    thread arrived first;
//  * labels: every triangle carries (parent, parity) = (a lower-numbered
//    triangle of its component, its flip relative to that triangle).  Each
//    iteration hooks to the lowest parent among the neighbours and then
//    jumps parent <- parent[parent]; both steps read one buffer and write
//    the other, so the result is independent of scheduling, and the labels
//    reach the component's lowest triangle in O(diameter) iterations at
//    most;
//  * non-orientability is checked once the labels are final: a link whose
//    two triangles' parities disagree with it;
//  * the signed volume is a fixed-order binary64 sum of v0 . (v1 x v2).
#include <algorithm>

#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

constexpr int kTBlock = 256;
constexpr unsigned long long kNoTri = ~0ull;

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void edge_clear_kernel(EdgeTable E) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E.size) return;
    E.key[i] = 0ull;
    E.count[i] = 0u;
    E.lo[i] = kNoTri;
    E.hi[i] = 0ull;
}

// Inserts the three edges of every triangle; slot[3 t + e] receives the
// table slot of edge e (u = idx[e], v = idx[e + 1 mod 3]) or -1 (u == v).
__global__ void edge_insert_kernel(const uint32_t* __restrict__ idx, int64_t nt, EdgeTable E,
                                   int64_t* __restrict__ slot) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint32_t v[3] = {idx[3 * t], idx[3 * t + 1], idx[3 * t + 2]};
    for (int e = 0; e < 3; ++e) {
        const uint32_t a = v[e], b = v[(e + 1) % 3];
        if (a == b) {
            slot[3 * t + e] = -1;
            continue;
        }
        const uint32_t mn = a < b ? a : b, mx = a < b ? b : a;
        const unsigned long long key = ((unsigned long long)mn << 32) | mx;  // >= 1: mx > mn >= 0
        const unsigned long long tag = ((unsigned long long)t << 1) | (a < b ? 0ull : 1ull);
        unsigned long long s = mix64(key) & (unsigned long long)(E.size - 1);
        // the table holds at least twice as many slots as edges, so a free
        // or matching slot is always found within E.size probes
        for (int64_t probe = 0; probe < E.size; ++probe) {
            const unsigned long long prev = atomicCAS(E.key + s, 0ull, key);
            if (prev == 0ull || prev == key) break;
            s = (s + 1) & (unsigned long long)(E.size - 1);
        }
        atomicAdd(E.count + s, 1u);
        atomicMin(E.lo + s, tag);
        atomicMax(E.hi + s, tag);
        slot[3 * t + e] = (int64_t)s;
    }
}

// nb[3 t + e]: the other triangle on edge e when exactly two triangles share
// it (else -1); same[3 t + e] = 1 when both traverse it the same way.
__global__ void edge_links_kernel(int64_t nt, EdgeTable E, const int64_t* __restrict__ slot, int32_t* __restrict__ nb,
                                  uint8_t* __restrict__ same, int* __restrict__ inconsistent) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nt) return;
    const int64_t t = i / 3;
    const int64_t s = slot[i];
    int32_t other = -1;
    uint8_t sm = 0;
    if (s >= 0 && E.count[s] == 2u) {
        const unsigned long long lo = E.lo[s], hi = E.hi[s];
        const int64_t a = (int64_t)(lo >> 1), b = (int64_t)(hi >> 1);
        if (a != b) {
            other = (int32_t)(a == t ? b : a);
            sm = ((lo ^ hi) & 1ull) == 0ull ? 1 : 0;
        }
    }
    nb[i] = other;
    same[i] = sm;
    if (sm) atomicOr(inconsistent, 1);
}

__global__ void label_init_kernel(int64_t nt, int32_t* __restrict__ parent, uint8_t* __restrict__ parity) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    parent[t] = (int32_t)t;
    parity[t] = 0;
}

// Hook: the lowest parent among t and its linked neighbours (first one on a
// tie), with the parity that the link implies.
__global__ void label_hook_kernel(int64_t nt, const int32_t* __restrict__ nb, const uint8_t* __restrict__ same,
                                  const int32_t* __restrict__ p_in, const uint8_t* __restrict__ q_in,
                                  int32_t* __restrict__ p_out, uint8_t* __restrict__ q_out, int* __restrict__ changed) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int32_t best = p_in[t];
    uint8_t bq = q_in[t];
    for (int e = 0; e < 3; ++e) {
        const int32_t n = nb[3 * t + e];
        if (n < 0) continue;
        const int32_t pn = p_in[n];
        if (pn < best) {
            best = pn;
            bq = q_in[n] ^ same[3 * t + e];
        }
    }
    p_out[t] = best;
    q_out[t] = bq;
    if (best != p_in[t]) atomicOr(changed, 1);
}

// Jump: parent <- parent[parent], parity composed along the way.
__global__ void label_jump_kernel(int64_t nt, const int32_t* __restrict__ p_in, const uint8_t* __restrict__ q_in,
                                  int32_t* __restrict__ p_out, uint8_t* __restrict__ q_out, int* __restrict__ changed) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int32_t p = p_in[t];
    const int32_t pp = p_in[p];
    p_out[t] = pp;
    q_out[t] = q_in[t] ^ q_in[p];
    if (pp != p) atomicOr(changed, 1);
}

// With final labels: a link whose parities disagree with it.
__global__ void orient_check_kernel(int64_t nt, const int32_t* __restrict__ nb, const uint8_t* __restrict__ same,
                                    const uint8_t* __restrict__ parity, int* __restrict__ non_orientable) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nt) return;
    const int32_t n = nb[i];
    if (n < 0) return;
    if (parity[i / 3] != (uint8_t)(parity[n] ^ same[i])) atomicOr(non_orientable, 1);
}

// Swaps the second and third index of every triangle with flip[t] != 0 (all
// triangles when flip is null).
__global__ void flip_kernel(uint32_t* __restrict__ idx, int64_t nt, const uint8_t* __restrict__ flip) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    if (flip && !flip[t]) return;
    const uint32_t a = idx[3 * t + 1];
    idx[3 * t + 1] = idx[3 * t + 2];
    idx[3 * t + 2] = a;
}

// Per-block binary64 partial sums of v0 . (v1 x v2) (6 x the signed volume).
__global__ __launch_bounds__(kTBlock) void volume_partial_kernel(const float* __restrict__ v,
                                                                 const uint32_t* __restrict__ idx, int64_t nt,
                                                                 double* __restrict__ part) {
    __shared__ double s[kTBlock];
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * kTBlock + threadIdx.x; t < nt; t += (int64_t)gridDim.x * kTBlock) {
        const int64_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
        const double ax = v[3 * a], ay = v[3 * a + 1], az = v[3 * a + 2];
        const double bx = v[3 * b], by = v[3 * b + 1], bz = v[3 * b + 2];
        const double cx = v[3 * c], cy = v[3 * c + 1], cz = v[3 * c + 2];
        acc += ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kTBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kTBlock - 1) / kTBlock); }

}  // namespace

int64_t edge_table_size(int64_t nt) {
    int64_t s = 1024;
    while (s < 6 * nt) s <<= 1;  // at least twice the 3 nt edges
    return s;
}

hipError_t launch_edge_links(const uint32_t* idx, int64_t nt, const EdgeTable& E, int64_t* slot, int32_t* nb,
                             uint8_t* same, int* inconsistent, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    edge_clear_kernel<<<blocks_for(E.size), kTBlock, 0, s>>>(E);
    edge_insert_kernel<<<blocks_for(nt), kTBlock, 0, s>>>(idx, nt, E, slot);
    edge_links_kernel<<<blocks_for(3 * nt), kTBlock, 0, s>>>(nt, E, slot, nb, same, inconsistent);
    return hipGetLastError();
}

hipError_t launch_label_init(int64_t nt, int32_t* parent, uint8_t* parity, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    label_init_kernel<<<blocks_for(nt), kTBlock, 0, s>>>(nt, parent, parity);
    return hipGetLastError();
}

hipError_t launch_label_step(int64_t nt, const int32_t* nb, const uint8_t* same, int32_t* p0, uint8_t* q0,
                             int32_t* p1, uint8_t* q1, int* changed, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    label_hook_kernel<<<blocks_for(nt), kTBlock, 0, s>>>(nt, nb, same, p0, q0, p1, q1, changed);
    label_jump_kernel<<<blocks_for(nt), kTBlock, 0, s>>>(nt, p1, q1, p0, q0, changed);
    return hipGetLastError();
}

hipError_t launch_orient_check(int64_t nt, const int32_t* nb, const uint8_t* same, const uint8_t* parity,
                               int* non_orientable, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    orient_check_kernel<<<blocks_for(3 * nt), kTBlock, 0, s>>>(nt, nb, same, parity, non_orientable);
    return hipGetLastError();
}

hipError_t launch_flip(uint32_t* idx, int64_t nt, const uint8_t* flip, hipStream_t s) {
    if (nt <= 0) return hipSuccess;
    flip_kernel<<<blocks_for(nt), kTBlock, 0, s>>>(idx, nt, flip);
    return hipGetLastError();
}

int volume_blocks(int64_t nt) { return (int)std::min<int64_t>(1024, std::max<int64_t>(1, (nt + kTBlock - 1) / kTBlock)); }

hipError_t launch_volume_partials(const float* v, const uint32_t* idx, int64_t nt, double* part, hipStream_t s) {
    volume_partial_kernel<<<(unsigned)volume_blocks(nt), kTBlock, 0, s>>>(v, idx, nt, part);
    return hipGetLastError();
}

}  // namespace rhz
