// Device FP32 vector/quaternion algebra for gfx950.
//
// Every expression keeps the reference's operation order so that results are
// bit-identical to the CPU engine (compiled with -ffp-contract=off, IEEE
// division and sqrt, no fast-math):
//   Quaternion::operator*  include/phys/math/quaternion.hpp:24-31
//   Quaternion::rotate     quaternion.hpp:47-51 (full q (v,0) q* sandwich)
//   Quaternion::toMat3     quaternion.hpp:54-64
//   Math3d::dot/length/normalize/cross/clamp  math_utils.cpp:7-31, math_utils.hpp:18-39
// std::max/std::min are written as (a < b) ? b : a and (b < a) ? b : a so the
// first argument wins ties exactly as in the reference (v_max_f32 would not).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {

struct f3 {
    float x, y, z;
};

struct quat {
    float x, y, z, w;
};

#define PG_D __device__ __forceinline__

PG_D f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PG_D f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PG_D f3 neg(f3 a) { return {-a.x, -a.y, -a.z}; }
PG_D f3 scale(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
PG_D f3 divide(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
PG_D float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PG_D f3 cross(f3 a, f3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PG_D float length(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
PG_D f3 normalize(f3 a)
{
    float len = length(a);
    if (len <= 1e-8f)
        return {0.0f, 0.0f, 0.0f};
    return scale(a, 1.0f / len);
}
PG_D float max_ref(float a, float b) { return (a < b) ? b : a; }
PG_D float min_ref(float a, float b) { return (b < a) ? b : a; }

PG_D quat qmul(quat a, quat b)
{
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
PG_D quat conj(quat q) { return {-q.x, -q.y, -q.z, q.w}; }
PG_D f3 rotate(quat q, f3 v)
{
    quat r = qmul(qmul(q, quat{v.x, v.y, v.z, 0.0f}), conj(q));
    return {r.x, r.y, r.z};
}

struct m3 {
    float m00, m01, m02, m10, m11, m12, m20, m21, m22;
};

PG_D m3 to_mat(quat q)
{
    float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    return {1.0f - 2.0f * (yy + zz), 2.0f * (xy - wz),        2.0f * (xz + wy),
            2.0f * (xy + wz),        1.0f - 2.0f * (xx + zz), 2.0f * (yz - wx),
            2.0f * (xz - wy),        2.0f * (yz + wx),        1.0f - 2.0f * (xx + yy)};
}
PG_D f3 col0(const m3& m) { return {m.m00, m.m10, m.m20}; }
PG_D f3 col1(const m3& m) { return {m.m01, m.m11, m.m21}; }
PG_D f3 col2(const m3& m) { return {m.m02, m.m12, m.m22}; }

struct pose {
    f3 p;
    quat q;
};

// narrowphase.cpp:16-25.
PG_D f3 to_local(f3 world, const pose& t) { return rotate(conj(t.q), sub(world, t.p)); }
PG_D f3 to_world(f3 local, const pose& t) { return add(t.p, rotate(t.q, local)); }

// Orderable bits of a finite float: monotone in the value, and -0 maps to the
// same key as +0 (the reference compares min.x with != and <, which treat
// them as equal, then falls back to the index).
PG_D uint32_t float_key(float f)
{
    uint32_t u = __float_as_uint(f);
    if (u == 0x80000000u)
        u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Resident body layout: pose = 2 float4 per body (p.xyz | 0, q.xyzw); shape
// = 2 float4 per body (data.xyz | type bits, hull first bits | hull count
// bits | 0 | 0).  A body's fields cost one or two cache lines.
PG_D pose load_pose4(const float4* __restrict__ p, uint32_t i)
{
    const float4 a = p[2 * size_t(i)], b = p[2 * size_t(i) + 1];
    return {{a.x, a.y, a.z}, {b.x, b.y, b.z, b.w}};
}
PG_D void store_pose4(float4* __restrict__ p, uint32_t i, const pose& t)
{
    p[2 * size_t(i)] = make_float4(t.p.x, t.p.y, t.p.z, 0.0f);
    p[2 * size_t(i) + 1] = make_float4(t.q.x, t.q.y, t.q.z, t.q.w);
}
PG_D f3 shape_data(const float4* __restrict__ s, uint32_t i)
{
    const float4 a = s[2 * size_t(i)];
    return {a.x, a.y, a.z};
}
PG_D uint32_t shape_type(const float4* __restrict__ s, uint32_t i)
{
    return __float_as_uint(s[2 * size_t(i)].w);
}
PG_D void hull_range(const float4* __restrict__ s, uint32_t i, uint32_t& first, uint32_t& count)
{
    const float4 b = s[2 * size_t(i) + 1];
    first = __float_as_uint(b.x);
    count = __float_as_uint(b.y);
}

} // namespace pg
