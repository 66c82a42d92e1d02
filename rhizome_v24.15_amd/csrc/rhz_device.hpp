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

// [F] one voxel of the sign apply (cuda_rasterizer.cu:650-661): a voxel still
// at the sentinel ships as 1e6, an inside voxel (sign 1) negated.
__device__ __forceinline__ float apply_sign(float d, uint8_t s) {
    if (d > kFarTest) d = 1e6f;
    return (s == 1) ? -d : d;
}
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

// The same closest-point regions in binary64 (the yardstick of the accuracy
// report; Ericson, Real-Time Collision Detection, 5.1.5).
__device__ __forceinline__ double point_tri_dist2_f64(double px, double py, double pz, const Tri9& t) {
    const double ax = t.ax, ay = t.ay, az = t.az;
    const double abx = (double)t.bx - ax, aby = (double)t.by - ay, abz = (double)t.bz - az;
    const double acx = (double)t.cx - ax, acy = (double)t.cy - ay, acz = (double)t.cz - az;
    const double apx = px - ax, apy = py - ay, apz = pz - az;
    const double d1 = abx * apx + aby * apy + abz * apz;
    const double d2 = acx * apx + acy * apy + acz * apz;
    if (d1 <= 0.0 && d2 <= 0.0) return apx * apx + apy * apy + apz * apz;
    const double bpx = px - t.bx, bpy = py - t.by, bpz = pz - t.bz;
    const double d3 = abx * bpx + aby * bpy + abz * bpz;
    const double d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= 0.0 && d4 <= d3) return bpx * bpx + bpy * bpy + bpz * bpz;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double v = d1 / (d1 - d3);
        const double qx = apx - v * abx, qy = apy - v * aby, qz = apz - v * abz;
        return qx * qx + qy * qy + qz * qz;
    }
    const double cpx = px - t.cx, cpy = py - t.cy, cpz = pz - t.cz;
    const double d5 = abx * cpx + aby * cpy + abz * cpz;
    const double d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= 0.0 && d5 <= d6) return cpx * cpx + cpy * cpy + cpz * cpz;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6);
        const double qx = apx - w * acx, qy = apy - w * acy, qz = apz - w * acz;
        return qx * qx + qy * qy + qz * qz;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const double qx = bpx - w * ((double)t.cx - t.bx);
        const double qy = bpy - w * ((double)t.cy - t.by);
        const double qz = bpz - w * ((double)t.cz - t.bz);
        return qx * qx + qy * qy + qz * qz;
    }
    const double den = 1.0 / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    const double qx = apx - abx * v - acx * w;
    const double qy = apy - aby * v - acy * w;
    const double qz = apz - abz * v - acz * w;
    return qx * qx + qy * qy + qz * qz;
}

// atan2 in FP32 for the winding sums: octant reduction to t = min/max in
// [0, 1], then atan(t) = t * P(t^2) with a degree-7 polynomial fitted to
// atan(t)/t on [0, 1] by iteratively reweighted least squares (Lawson's
// minimax method); max |error| 1.1e-7 rad evaluated in FP32.  Zero for
// (0, 0), signed by y.
__device__ __forceinline__ float atan2_poly(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = (mx > 0.0f) ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
    const float s = t * t;
    float p = -0.004055280238389969f;
    p = __builtin_fmaf(p, s, 0.021865447983145714f);
    p = __builtin_fmaf(p, s, -0.05591578036546707f);
    p = __builtin_fmaf(p, s, 0.0964243933558464f);
    p = __builtin_fmaf(p, s, -0.13908720016479492f);
    p = __builtin_fmaf(p, s, 0.1994658261537552f);
    p = __builtin_fmaf(p, s, -0.33329862356185913f);
    p = __builtin_fmaf(p, s, 0.9999993443489075f);
    float r = p * t;
    r = (ay > ax) ? (1.5707963267948966f - r) : r;
    r = (x < 0.0f) ? (3.141592653589793f - r) : r;
    return copysignf(r, y);
}

// Van Oosterom-Strackee solid angle in FP32 (reference triangle_solid_angle,
// math_utils.cuh:109-132: vertex guard <= 1e-6 returns 0).  The winding
// kernels use this with hardware sqrt and fused dot products; the sign
// decision is protected by the FP64 recompute near |w| = 0.5.
__device__ __forceinline__ float solid_angle_fast(float px, float py, float pz, const Tri9& t) {
    float ax = t.ax - px, ay = t.ay - py, az = t.az - pz;
    float bx = t.bx - px, by = t.by - py, bz = t.bz - pz;
    float cx = t.cx - px, cy = t.cy - py, cz = t.cz - pz;
    float la = __builtin_amdgcn_sqrtf(__builtin_fmaf(az, az, __builtin_fmaf(ay, ay, ax * ax)));
    float lb = __builtin_amdgcn_sqrtf(__builtin_fmaf(bz, bz, __builtin_fmaf(by, by, bx * bx)));
    float lc = __builtin_amdgcn_sqrtf(__builtin_fmaf(cz, cz, __builtin_fmaf(cy, cy, cx * cx)));
    if (la <= 1e-6f || lb <= 1e-6f || lc <= 1e-6f) return 0.0f;
    float crx = __builtin_fmaf(by, cz, -bz * cy);
    float cry = __builtin_fmaf(bz, cx, -bx * cz);
    float crz = __builtin_fmaf(bx, cy, -by * cx);
    float det = __builtin_fmaf(az, crz, __builtin_fmaf(ay, cry, ax * crx));
    float dab = __builtin_fmaf(az, bz, __builtin_fmaf(ay, by, ax * bx));
    float dbc = __builtin_fmaf(bz, cz, __builtin_fmaf(by, cy, bx * cx));
    float dca = __builtin_fmaf(cz, az, __builtin_fmaf(cy, ay, cx * ax));
    float denom = __builtin_fmaf(la * lb, lc, __builtin_fmaf(dab, lc, __builtin_fmaf(dbc, la, dca * lb)));
    return 2.0f * atan2_poly(det, denom);
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

// x / 3 correctly rounded for every finite x except -0 (which maps to +0):
// q0 = x * RN(1/3), exact residual r = x - 3 q0 by FMA, q = q0 + r * RN(1/3).
// Checked against IEEE division over all 2^32 bit patterns; the sweep
// numerators are sums of non-negative terms, never -0.
__device__ __forceinline__ float div3(float x) {
    const float c = 0.3333333432674408f;  // RN(1/3)
    float q0 = x * c;
    float r = __builtin_fmaf(-q0, 3.0f, x);
    return __builtin_fmaf(r, c, q0);
}

// The same update as godunov() (identical results bit for bit) arranged for
// the sweeps: constants 2h^2 and 3h^2 precomputed with the reference's
// evaluation order ((2h)h, (3h)h), an exit when no branch can go below `cur`
// (with all three minima finite every branch returns at least
// s0 + h/sqrt(3) > s0 + 0.57 h, so cur <= s0 + 0.57 h cannot be improved),
// and early exits before each square root.
// Returns min(cur, G).
__device__ __forceinline__ float godunov_min(float cur, float ax, float ay, float az, float h, float h2x2,
                                             float h2x3, float h057, float hbig) {
    float s0 = fminf(ax, fminf(ay, az));
    float s2 = fmaxf(ax, fmaxf(ay, az));
    // The bound needs s0 <= s1 <= s2, which the sum-minus-extremes middle value
    // only delivers when no axis holds the 1e30 sentinel; below 4096 h the
    // rounding of s0 + 0.57 h stays far inside the margin.
    if (!(cur > s0 + h057) && s2 < hbig) return cur;
    float s1 = ax + ay + az - s0 - s2;
    float u = s0 + h;
    if (u > s1) {
        float d10 = s1 - s0;
        float disc = h2x2 - d10 * d10;
        if (disc >= 0.0f) {
            u = (s0 + s1 + sqrtf(disc)) * 0.5f;
            if (u > s2) {
                float d21 = s2 - s1, d20 = s2 - s0;
                float disc3 = h2x3 - (d10 * d10 + d21 * d21 + d20 * d20) * 0.5f;
                if (disc3 >= 0.0f) u = div3(s0 + s1 + s2 + sqrtf(disc3));
            }
        }
    }
    return (u < cur) ? u : cur;
}

__device__ __forceinline__ float lattice(int32_t mn, int i, float h) { return (float)(mn + i) * h; }

// Block order for 8-XCD parts: dispatch hands block b to XCD b % 8, so
// consecutive logical tiles are mapped to the same XCD (and its L2).
__device__ __forceinline__ unsigned xcd_logical_block(unsigned b, unsigned nblocks) {
    const unsigned per = nblocks / 8u;
    const unsigned full = per * 8u;
    if (b >= full) return b;  // tail blocks keep their index
    return (b % 8u) * per + b / 8u;
}

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
