// rhz_device.hpp -- device math shared by the gfx950 kernels.
//
// All translation units are compiled with -ffp-contract=off: each a*b+c below
// rounds twice, so band distances, scanline hit heights and Godunov updates are
// the IEEE binary32 values of the written expressions.  Kernels that want fused
// multiply-adds (the winding sums) call __builtin_fmaf explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rhz {

constexpr float kFar = 1e30f;            // unreached marker (cuda_rasterizer.cu:77)
constexpr float kFarTest = 1e29f;        // audit threshold (:531)
constexpr float kInv4Pi = 0.07957747154594767f;
constexpr int kWave = 64;

struct Tri9 {
    float ax, ay, az, bx, by, bz, cx, cy, cz;
};

// Triangles are 12 floats (v0, v1, v2, normal); three float4 loads.
__device__ __forceinline__ Tri9 load_tri(const float* __restrict__ t, int64_t i) {
    const float4* p = reinterpret_cast<const float4*>(t + 12 * i);
    float4 a = p[0], b = p[1], c = p[2];
    return Tri9{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x};
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;  // (x + y) + z, no contraction
}

// Squared distance from p to triangle (a, b, c): Ericson's closest-point
// regions, same branch structure and operation order as the reference
// point_triangle_distance_sq (native/src/core/math_utils.cuh:43-106).
__device__ __forceinline__ float point_tri_dist2(float px, float py, float pz, const Tri9& t) {
    float abx = t.bx - t.ax, aby = t.by - t.ay, abz = t.bz - t.az;
    float acx = t.cx - t.ax, acy = t.cy - t.ay, acz = t.cz - t.az;
    float apx = px - t.ax, apy = py - t.ay, apz = pz - t.az;
    float d1 = dot3(abx, aby, abz, apx, apy, apz);
    float d2 = dot3(acx, acy, acz, apx, apy, apz);
    if (d1 <= 0.0f && d2 <= 0.0f) return dot3(apx, apy, apz, apx, apy, apz);

    float bpx = px - t.bx, bpy = py - t.by, bpz = pz - t.bz;
    float d3 = dot3(abx, aby, abz, bpx, bpy, bpz);
    float d4 = dot3(acx, acy, acz, bpx, bpy, bpz);
    if (d3 >= 0.0f && d4 <= d3) return dot3(bpx, bpy, bpz, bpx, bpy, bpz);

    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        float qx = px - (t.ax + v * abx), qy = py - (t.ay + v * aby), qz = pz - (t.az + v * abz);
        return dot3(qx, qy, qz, qx, qy, qz);
    }

    float cpx = px - t.cx, cpy = py - t.cy, cpz = pz - t.cz;
    float d5 = dot3(abx, aby, abz, cpx, cpy, cpz);
    float d6 = dot3(acx, acy, acz, cpx, cpy, cpz);
    if (d6 >= 0.0f && d5 <= d6) return dot3(cpx, cpy, cpz, cpx, cpy, cpz);

    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float w = d2 / (d2 - d6);
        float qx = px - (t.ax + w * acx), qy = py - (t.ay + w * acy), qz = pz - (t.az + w * acz);
        return dot3(qx, qy, qz, qx, qy, qz);
    }

    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        float qx = px - (t.bx + w * (t.cx - t.bx));
        float qy = py - (t.by + w * (t.cy - t.by));
        float qz = pz - (t.bz + w * (t.cz - t.bz));
        return dot3(qx, qy, qz, qx, qy, qz);
    }

    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom;
    float w = vc * denom;
    float qx = px - (t.ax + abx * v + acx * w);
    float qy = py - (t.ay + aby * v + acy * w);
    float qz = pz - (t.az + abz * v + acz * w);
    return dot3(qx, qy, qz, qx, qy, qz);
}

// Van Oosterom-Strackee solid angle in FP32 (reference triangle_solid_angle,
// math_utils.cuh:109-132: vertex guard <= 1e-6 returns 0).  The winding
// kernels use this with hardware sqrt and fused dot products; the sign
// decision is protected by the FP64 recompute near |w| = 0.5.
__device__ __forceinline__ float solid_angle_fast(float px, float py, float pz, const Tri9& t) {
    float ax = t.ax - px, ay = t.ay - py, az = t.az - pz;
    float bx = t.bx - px, by = t.by - py, bz = t.bz - pz;
    float cx = t.cx - px, cy = t.cy - py, cz = t.cz - pz;
    float la = __builtin_sqrtf(__builtin_fmaf(az, az, __builtin_fmaf(ay, ay, ax * ax)));
    float lb = __builtin_sqrtf(__builtin_fmaf(bz, bz, __builtin_fmaf(by, by, bx * bx)));
    float lc = __builtin_sqrtf(__builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx)));
    if (la <= 1e-6f || lb <= 1e-6f || lc <= 1e-6f) return 0.0f;
    float crx = __builtin_fmaf(by, cz, -bz * cy);
    float cry = __builtin_fmaf(bz, cx, -bx * cz);
    float crz = __builtin_fmaf(bx, cy, -by * cx);
    float det = __builtin_fmaf(az, crz, __builtin_fmaf(ay, cry, ax * crx));
    float dab = __builtin_fmaf(az, bz, __builtin_fmaf(ay, by, ax * bx));
    float dbc = __builtin_fmaf(bz, cz, __builtin_fmaf(by, cy, bx * cx));
    float dca = __builtin_fmaf(cz, az, __builtin_fmaf(cy, ay, cx * ax));
    float denom = __builtin_fmaf(la * lb, lc, __builtin_fmaf(dab, lc, __builtin_fmaf(dbc, la, dca * lb)));
    return 2.0f * atan2f(det, denom);
}

// Same quantity in binary64 (the tie-break recompute).
__device__ __forceinline__ double solid_angle_f64(double px, double py, double pz, const Tri9& t) {
    double ax = (double)t.ax - px, ay = (double)t.ay - py, az = (double)t.az - pz;
    double bx = (double)t.bx - px, by = (double)t.by - py, bz = (double)t.bz - pz;
    double cx = (double)t.cx - px, cy = (double)t.cy - py, cz = (double)t.cz - pz;
    double la = sqrt(ax * ax + ay * ay + az * az);
    double lb = sqrt(bx * bx + by * by + bz * bz);
    double lc = sqrt(cx * cx + cy * cy + cz * cz);
    if (la <= 1e-6 || lb <= 1e-6 || lc <= 1e-6) return 0.0;
    double det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
    double denom = la * lb * lc + (ax * bx + ay * by + az * bz) * lc +
                   (bx * cx + by * cy + bz * cz) * la + (cx * ax + cy * ay + cz * az) * lb;
    return 2.0 * atan2(det, denom);
}

// Reference-order FP32 solid angle (no contraction, correctly rounded sqrt),
// used by the plain-FP32 precision mode.
__device__ __forceinline__ float solid_angle_ref(float px, float py, float pz, const Tri9& t) {
    float ax = t.ax - px, ay = t.ay - py, az = t.az - pz;
    float bx = t.bx - px, by = t.by - py, bz = t.bz - pz;
    float cx = t.cx - px, cy = t.cy - py, cz = t.cz - pz;
    float la = sqrtf(dot3(ax, ay, az, ax, ay, az));
    float lb = sqrtf(dot3(bx, by, bz, bx, by, bz));
    float lc = sqrtf(dot3(cx, cy, cz, cx, cy, cz));
    if (la <= 1e-6f || lb <= 1e-6f || lc <= 1e-6f) return 0.0f;
    float det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
    float denom = la * lb * lc + dot3(ax, ay, az, bx, by, bz) * lc + dot3(bx, by, bz, cx, cy, cz) * la +
                  dot3(cx, cy, cz, ax, ay, az) * lb;
    return 2.0f * atan2f(det, denom);
}

// Godunov upwind update from the per-axis neighbour minima; operation order of
// the reference rhi_eikonal_godunov (cuda_rasterizer.cu:405-425), including the
// sum-minus-extremes middle value and the halved 3-D discriminant.
__device__ __forceinline__ float godunov(float ax, float ay, float az, float h) {
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
                              ((s1 - s0) * (s1 - s0) + (s2 - s1) * (s2 - s1) + (s2 - s0) * (s2 - s0)) * 0.5f;
                if (disc3 >= 0.0f) updated = (s0 + s1 + s2 + sqrtf(disc3)) / 3.0f;
            }
        }
    }
    return updated;
}

__device__ __forceinline__ float lattice(int32_t mn, int i, float h) { return (float)(mn + i) * h; }

// Wave-aggregated append of `want` items; returns this lane's slot or -1.
__device__ __forceinline__ long long wave_append(unsigned long long* counter, bool want) {
    unsigned long long mask = __ballot(want);
    if (mask == 0) return -1;
    int lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    if (!want) return -1;
    unsigned long long below = mask & ((1ull << lane) - 1ull);
    return (long long)(base + __popcll(below));
}

}  // namespace rhz
