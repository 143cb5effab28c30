// CPU restatement of the phys-engine collision path — TEST INFRASTRUCTURE.
//
// This is the oracle: the checker the HIP path is compared against, and the
// "port" CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it (as liboracle.so).  The product never links it.
//
// Every function replays the reference's FP32 arithmetic operation by
// operation (compiled with -ffp-contract=off); the scalar type is a template
// parameter so the same code also gives the FP64 re-evaluation.
//
// Parity pins (see tests/test_oracle_*.py): each function below is checked
// against oracle/_ref (the reference compiled from its own sources) on the
// committed golden fixtures, and against closed forms and reference KATs.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace orc {

template <typename T>
struct V3 {
    T x, y, z;
};

template <typename T>
inline V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T>
inline V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T>
inline V3<T> neg(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T>
inline V3<T> scale(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T>
inline V3<T> divide(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
// Math3d::dot, math_utils.cpp:25-28 (left-to-right sum).
template <typename T>
inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Math3d::cross, math_utils.hpp:18-25.
template <typename T>
inline V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// |x| and sqrt as the reference calls them; overloaded for the dual numbers
// of the FP64 adjoint.
template <typename T>
inline T absOf(T x) { return std::abs(x); }
template <typename T>
inline T sqrtOf(T x) { return std::sqrt(x); }
// Largest finite value (std::numeric_limits<T>::max()).
template <typename T>
inline T bigOf() { return std::numeric_limits<T>::max(); }

// Math3d::length, math_utils.cpp:7-12.
template <typename T>
inline T length(V3<T> a) { return sqrtOf(a.x * a.x + a.y * a.y + a.z * a.z); }
// Math3d::normalize, math_utils.cpp:19-23: zero below 1e-8, else v * (1/len).
template <typename T>
inline V3<T> normalize(V3<T> a)
{
    T len = length(a);
    if (len <= T(1e-8f))
        return {T(0), T(0), T(0)};
    return scale(a, T(1) / len);
}
// std::max / std::min as the reference calls them (first argument on ties).
template <typename T>
inline T maxOf(T a, T b) { return (a < b) ? b : a; }
template <typename T>
inline T minOf(T a, T b) { return (b < a) ? b : a; }

template <typename T>
struct Quat {
    T x, y, z, w;
};

// Quaternion::operator*, quaternion.hpp:24-31.
template <typename T>
inline Quat<T> qmul(Quat<T> a, Quat<T> b)
{
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <typename T>
inline Quat<T> conj(Quat<T> q) { return {-q.x, -q.y, -q.z, q.w}; }
// Quaternion::rotate, quaternion.hpp:47-51: the full sandwich q * (v,0) * q*.
template <typename T>
inline V3<T> rotate(Quat<T> q, V3<T> v)
{
    Quat<T> r = qmul(qmul(q, Quat<T>{v.x, v.y, v.z, T(0)}), conj(q));
    return {r.x, r.y, r.z};
}

template <typename T>
struct M3 {
    T m[3][3];
};
// Quaternion::toMat3, quaternion.hpp:54-64.
template <typename T>
inline M3<T> toMat(Quat<T> q)
{
    T xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    T xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    T wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    M3<T> r;
    r.m[0][0] = T(1) - T(2) * (yy + zz);
    r.m[0][1] = T(2) * (xy - wz);
    r.m[0][2] = T(2) * (xz + wy);
    r.m[1][0] = T(2) * (xy + wz);
    r.m[1][1] = T(1) - T(2) * (xx + zz);
    r.m[1][2] = T(2) * (yz - wx);
    r.m[2][0] = T(2) * (xz - wy);
    r.m[2][1] = T(2) * (yz + wx);
    r.m[2][2] = T(1) - T(2) * (xx + yy);
    return r;
}
template <typename T>
inline V3<T> column(const M3<T>& m, int c) { return {m.m[0][c], m.m[1][c], m.m[2][c]}; }

template <typename T>
struct Pose {
    V3<T> p;
    Quat<T> q;
};

// narrowphase.cpp:16-25.
template <typename T>
inline V3<T> toLocal(V3<T> world, const Pose<T>& t) { return rotate(conj(t.q), sub(world, t.p)); }
template <typename T>
inline V3<T> toWorld(V3<T> local, const Pose<T>& t) { return add(t.p, rotate(t.q, local)); }

enum ShapeType : uint8_t { kSphere = 0, kBox = 1, kHull = 2 };

template <typename T>
struct Shape {
    int type;
    V3<T> data;              // sphere radius in x; box half extents
    const V3<T>* points;     // hull points (collider local)
    uint32_t pointCount;
};

template <typename T>
struct Contact {
    V3<T> normal;
    uint32_t count;
    V3<T> anchorA[4];
    V3<T> anchorB[4];
    T depth[4];
    V3<T> world[4];  // world contact points (anchors are their collider-local forms)
};

// Discrete choices of one sphere-box or box-box contact: equal choices at two
// poses mean the contact is one smooth function between them.
struct Choice {
    int sel = -1, flip = 0, fallback = 0, count = 0;  // box-box SAT axis, sign, fallback
    int side[4] = {0, 0, 0, 0}, corner[4] = {0, 0, 0, 0}, zero[4] = {0, 0, 0, 0};
    // Box-box: sign bits of the other box's support corner per side.  Not part
    // of the comparison: when the support is a face or an edge of the other
    // box, every tied corner has the same projection on the normal for all
    // nearby poses (the normal is that face's or edge's own axis).
    int support[2] = {0, 0};
    int branch = 0, face = 0, clamped = 0;             // sphere-box
    bool operator==(const Choice& o) const
    {
        if (sel != o.sel || flip != o.flip || fallback != o.fallback || count != o.count ||
            branch != o.branch || face != o.face || clamped != o.clamped)
            return false;
        for (int k = 0; k < 4; ++k)
            if (side[k] != o.side[k] || corner[k] != o.corner[k] || zero[k] != o.zero[k])
                return false;
        return true;
    }
};

struct NarrowOptions {
    uint32_t gjkIterations = 64;
    uint32_t epaIterations = 64;
    float epaTolerance = 1e-6f;
    // Bounded polytope (0 = unbounded, the reference's behaviour): a pair whose
    // EPA would hold more faces, or whose horizon would grow past maxEdges,
    // stops with an overflow status, exactly where the HIP kernel stops.
    uint32_t maxFaces = 0;
    uint32_t maxEdges = 0;
    // Past those bounds, continue in multiplicity form (epaContinue) with at
    // most multiFaces distinct faces (0: stop with the overflow status).
    uint32_t multiFaces = 0;
    bool multiExactOrder = true;
};

// Narrow phase entry (NarrowPhase::generateContact, narrowphase.cpp:326-374,
// plus GJK/EPA for hull pairs).  Returns true on contact; `c` is reset.
template <typename T>
bool generateContact(const Shape<T>& a, const Pose<T>& ta, const Shape<T>& b, const Pose<T>& tb,
                     const NarrowOptions& opt, Contact<T>& c, int* gjkStatus = nullptr);

} // namespace orc
