// spt_math.h - float3 arithmetic shared by host set-up code and gfx950 kernels.
//
// Every formula follows the reference's GLM 0.9.9 scalar forms so results are
// bit-identical on host and device (compiled with -ffp-contract=off, IEEE
// divide/sqrt): dot = (x*x' + y*y') + z*z'; cross(a, b) = (a.y*b.z - b.y*a.z,
// a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y); normalize(v) = v * (1 / sqrt(dot(v, v)));
// min(a, b) = (b < a) ? b : a; max(a, b) = (a < b) ? b : a;
// mat3 * v = col0*v.x + col1*v.y + col2*v.z.
#pragma once

#include <stdint.h>

#include <cmath>

#include "spt_fmath.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SPT_HD __host__ __device__ inline
#else
#define SPT_HD inline
#endif

namespace spt {

static constexpr float kPi = 3.141592653589793238462643383279502884f;  // core/core.h:4

struct V3 {
    float x, y, z;
};
SPT_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
SPT_HD V3 v3s(float s) { return V3{s, s, s}; }
SPT_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SPT_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SPT_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
SPT_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
SPT_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
SPT_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
SPT_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
SPT_HD V3 operator-(V3 a, float s) { return V3{a.x - s, a.y - s, a.z - s}; }
SPT_HD V3 operator+(V3 a, float s) { return V3{a.x + s, a.y + s, a.z + s}; }
SPT_HD V3 operator-(float s, V3 a) { return V3{s - a.x, s - a.y, s - a.z}; }
SPT_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SPT_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
SPT_HD float gmin(float a, float b) { return (b < a) ? b : a; }
SPT_HD float gmax(float a, float b) { return (a < b) ? b : a; }
SPT_HD V3 vmin(V3 a, V3 b) { return V3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
SPT_HD V3 vmax(V3 a, V3 b) { return V3{gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }

// Correctly rounded sqrt. On gfx950 __builtin_sqrtf lowers to v_sqrt_f32 plus
// an fma-based one-ulp correction; __fsqrt_rn lowers to the bare ~1-ulp
// v_sqrt_f32 and must not be used here.
SPT_HD float sqrt_rn(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_sqrtf(x);
#else
    return std::sqrt(x);
#endif
}
SPT_HD V3 normalize(V3 v) { return v * (1.0f / sqrt_rn(dot(v, v))); }

// 3x3 basis by columns; l2w = (u, v, w), w2l = its transpose.
struct Frame {
    V3 u, v, w;
};
SPT_HD V3 to_world(const Frame& f, V3 a) {
    return V3{f.u.x * a.x + f.v.x * a.y + f.w.x * a.z, f.u.y * a.x + f.v.y * a.y + f.w.y * a.z,
              f.u.z * a.x + f.v.z * a.y + f.w.z * a.z};
}
SPT_HD V3 to_local(const Frame& f, V3 a) { return V3{dot(f.u, a), dot(f.v, a), dot(f.w, a)}; }

// ONB(v), core/onb.cpp:3-16
SPT_HD Frame frame_from_up(V3 v) {
    V3 w;
    if (std::fabs(v.x) > std::fabs(v.y))
        w = normalize(V3{v.z, 0.0f, -v.x});
    else
        w = normalize(V3{0.0f, -v.z, v.y});
    return Frame{cross(v, w), v, w};
}
// ONB(u, w), core/onb.cpp:18-27
SPT_HD Frame frame_from_uw(V3 u, V3 w) { return Frame{u, cross(w, u), w}; }

// AABB::getArea, acceleration/aabb.cpp:11-16
SPT_HD float box_area(V3 mn, V3 mx) {
    float l = (mx.x - mn.x);
    float m = (mx.y - mn.y);
    float n = (mx.z - mn.z);
    return 2.0f * (l * (m + n) + m * n);
}

// Slab test with the reference's exact compare/select structure
// (acceleration/aabb.cpp:31-99; Williams et al. 2005). inv = 1/d per component
// (IEEE divide). Returns the reference's boolean; when true, [tn, tf] is the
// parametric interval of the ray's LINE inside the box (no t >= 0 clamp).
// NaN from 0 * inf falls through the comparisons exactly as on the host.
// The early exits of the reference are evaluated without branching: every
// value is computed and the two pass conditions are combined at the end,
// which gives the same boolean (nothing computed has side effects) and lets a
// GPU issue all box loads before the first comparison.
SPT_HD bool slab(V3 o, V3 inv, V3 mn, V3 mx, float& tn, float& tf) {
    const bool px = inv.x >= 0.0f, py = inv.y >= 0.0f, pz = inv.z >= 0.0f;
    float tmin = ((px ? mn.x : mx.x) - o.x) * inv.x;
    float tmax = ((px ? mx.x : mn.x) - o.x) * inv.x;
    const float tymin = ((py ? mn.y : mx.y) - o.y) * inv.y;
    const float tymax = ((py ? mx.y : mn.y) - o.y) * inv.y;
    const float tzmin = ((pz ? mn.z : mx.z) - o.z) * inv.z;
    const float tzmax = ((pz ? mx.z : mn.z) - o.z) * inv.z;
    const bool pass_xy = !((tmin > tymax) || (tymin > tmax));
    tmin = (tymin > tmin) ? tymin : tmin;
    tmax = (tymax < tmax) ? tymax : tmax;
    const bool pass_z = !((tmin > tzmax) || (tzmin > tmax));
    tn = (tzmin > tmin) ? tzmin : tmin;
    tf = (tzmax < tmax) ? tzmax : tmax;
    return pass_xy & pass_z;
}

// AABB::isPointInside, acceleration/aabb.cpp:101-105
SPT_HD bool point_inside(V3 p, V3 mn, V3 mx) {
    const float e = 1e-3f;
    return (p.x >= (mn.x - e)) && (p.x <= (mx.x + e)) && (p.y >= (mn.y - e)) && (p.y <= (mx.y + e)) &&
           (p.z >= (mn.z - e)) && (p.z <= (mx.z + e));
}

// Moller-Trumbore, non-culling branch, primitives/triangle.cpp:22-65 with the
// edges precomputed exactly as the reference computes them per test. Every
// quantity is computed and the reference's rejections (|det| < 1e-6, u and v
// outside the triangle) are combined at the end: the same acceptance and the
// same t, u, v, without data-dependent branches.
SPT_HD bool tri_test(V3 o, V3 d, V3 v0, V3 e1, V3 e2, float& t, float& u, float& v) {
    const V3 pvec = cross(d, e2);
    const float det = dot(e1, pvec);
    const bool det_ok = !((det > -1e-6f) && (det < 1e-6f));
    const float inv_det = 1.0f / det;
    const V3 tvec = o - v0;
    u = dot(tvec, pvec) * inv_det;
    const bool u_ok = !((u < 0.0f) || (u > 1.0f));
    const V3 qvec = cross(tvec, e1);
    v = dot(d, qvec) * inv_det;
    const bool v_ok = !((v < 0.0f) || ((u + v) > 1.0f));
    t = dot(e2, qvec) * inv_det;
    return det_ok & u_ok & v_ok;
}

// Sphere::intersect (primitives/sphere.cpp:6-38, Hultquist 1990): a hit
// whenever the line meets the sphere; t is the near root when it exceeds
// 1e-5, else the far one (possibly negative: the caller's t > 0 test
// rejects it).
SPT_HD bool sphere_test(V3 o, V3 d, V3 c, float r, float& t) {
    const V3 eo = c - o;
    const float v = dot(eo, d);
    const float disc = (r * r) - (dot(eo, eo) - (v * v));
    const float sq = sqrt_rn(disc);  // NaN when disc < 0, rejected below
    const float t1 = v - sq, t2 = v + sq;
    t = (t1 > 1e-5f) ? t1 : t2;
    return !(disc < 0.0f);
}

// Counter-based draws (definition in include/spt.h).
static constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
SPT_HD uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
SPT_HD uint64_t rng_key(uint64_t seed, uint64_t path) { return mix64(seed ^ (path * kGolden)); }
SPT_HD float rng_u(uint64_t key, uint32_t k) {
    uint64_t z = mix64(key + (uint64_t)(k + 1) * kGolden);
    return (float)(z >> 40) * (1.0f / 16777216.0f);
}

// Float transcendentals: float-only routines that give the same bits on the
// host and the device (spt_fmath.h); no FP64 on the device.
SPT_HD void f_sincos(float x, float& s, float& c) { fm_sincos(x, s, c); }
SPT_HD float f_exp(float x) { return fm_exp(x); }
SPT_HD float f_log(float x) { return fm_log(x); }
SPT_HD float f_atan(float x) { return fm_atan(x); }

}  // namespace spt
