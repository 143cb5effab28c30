// Forward contact generation for sphere-sphere, sphere-box and box-box
// (src/collision/narrowphase.cpp:16-324, sat.cpp:38-102), shared by the
// narrow-phase kernels and the adjoint kernel, which recomputes the forward
// with the same FP32 operations and therefore the same discrete choices.
// All arithmetic replays the reference's FP32 operation order (devmath.hpp).
#pragma once

#include "devmath.hpp"
#include "internal.hpp"

namespace pg {
namespace {

struct Bodies {
    uint32_t n;
    const float4* pose;   // 2 float4 per body (load_pose4)
    const float4* shape;  // 2 float4 per body (shape_data, shape_type, hull_range)
    const float* hpts;    // hull points, 3 planes of nh
    uint32_t nh;
};

struct World {
    const uint32_t* body;
    const float* local_pose;
    const float* friction;
    const float* restitution;
};

struct Contact {
    f3 normal;
    uint32_t count;
    f3 a[4];
    f3 b[4];
    float depth[4];
};

__device__ __forceinline__ pose load_pose(const Bodies& B, uint32_t i)
{
    return load_pose4(B.pose, i);
}
__device__ __forceinline__ f3 load_data(const Bodies& B, uint32_t i)
{
    return shape_data(B.shape, i);
}
__device__ __forceinline__ uint32_t load_type(const Bodies& B, uint32_t i)
{
    return shape_type(B.shape, i);
}

__device__ __forceinline__ uint32_t pair_class(uint32_t ta, uint32_t tb)
{
    const uint32_t lo = min(ta, tb), hi = max(ta, tb);
    return hi < 2 ? lo + hi : 3 + lo;
}

__device__ __forceinline__ void set_anchors(Contact& c, uint32_t k, f3 w, const pose& ta,
                                            const pose& tb)
{
    c.a[k] = to_local(w, ta);
    c.b[k] = to_local(w, tb);
}

// ------------------------------------------------------------ sphere-sphere
// intersectSphereSphere, narrowphase.cpp:132-160.
__device__ bool sphere_sphere(float ra, const pose& ta, float rb, const pose& tb, Contact& c)
{
    const f3 delta = sub(tb.p, ta.p);
    const float dist = length(delta);
    const float radii = ra + rb;
    if (dist >= radii)
        return false;
    c.normal = dist > 1e-6f ? divide(delta, dist) : f3{1.0f, 0.0f, 0.0f};
    c.count = 1;
    c.depth[0] = radii - dist;
    const f3 pa = add(ta.p, scale(c.normal, ra));
    const f3 pb = sub(tb.p, scale(c.normal, rb));
    set_anchors(c, 0, scale(add(pa, pb), 0.5f), ta, tb);
    return true;
}

// ------------------------------------------------------------ sphere-box
// Math3d::clamp (math_utils.hpp:31-39); min > max cannot happen for extents >= 0.
__device__ __forceinline__ float clamp_ref(float v, float lo, float hi)
{
    if (lo == hi)
        return lo;
    if (v < lo)
        return lo;
    if (v > hi)
        return hi;
    return v;
}

// intersectSphereBox, narrowphase.cpp:162-248.
__device__ bool sphere_box(float r, const pose& ts, f3 h, const pose& tb, Contact& c)
{
    const f3 cl = rotate(conj(tb.q), sub(ts.p, tb.p));
    const f3 closest{clamp_ref(cl.x, -h.x, h.x), clamp_ref(cl.y, -h.y, h.y),
                     clamp_ref(cl.z, -h.z, h.z)};
    const f3 dl = sub(closest, cl);
    const float dist = length(dl);
    f3 nl;
    float pen;
    if (dist > 1e-6f) {
        if (dist >= r)
            return false;
        nl = divide(dl, dist);
        pen = r - dist;
    } else {
        // Nearest face, strict <, order +X -X +Y -Y +Z -Z (:198-233).
        float nearest = h.x - cl.x;
        nl = {-1.0f, 0.0f, 0.0f};
        float d = h.x + cl.x;
        if (d < nearest) { nearest = d; nl = {1.0f, 0.0f, 0.0f}; }
        d = h.y - cl.y;
        if (d < nearest) { nearest = d; nl = {0.0f, -1.0f, 0.0f}; }
        d = h.y + cl.y;
        if (d < nearest) { nearest = d; nl = {0.0f, 1.0f, 0.0f}; }
        d = h.z - cl.z;
        if (d < nearest) { nearest = d; nl = {0.0f, 0.0f, -1.0f}; }
        d = h.z + cl.z;
        if (d < nearest) { nearest = d; nl = {0.0f, 0.0f, 1.0f}; }
        pen = r + nearest;
    }
    c.normal = rotate(tb.q, nl);
    c.count = 1;
    c.depth[0] = pen;
    const f3 box_point = to_world(closest, tb);
    const f3 sphere_point = add(ts.p, scale(c.normal, r));
    set_anchors(c, 0, scale(add(sphere_point, box_point), 0.5f), ts, tb);
    return true;
}

// ------------------------------------------------------------ box-box
// Discrete choices of one box-box forward, recorded for the adjoint: the SAT
// axis (0-2 A's axes, 3-5 B's, 6 + 3i + j = Ai x Bj), the normal flip, and for
// each contact point its source face (0: A along n, 1: B along -n), the face
// vertex, and whether its depth was clamped to 0; `fallback` when no face
// vertex was kept.
struct BoxTape {
    uint32_t sel;
    bool flip;
    bool fallback;
    uint32_t count;
    uint8_t side[4];
    uint8_t corner[4];
    bool zero[4];
};

// projectedRadius (sat.cpp:38-42) == boxProjectedRadius (narrowphase.cpp:37-46).
__device__ __forceinline__ float proj_radius(f3 h, const m3& m, f3 axis)
{
    return fabsf(dot(col0(m), axis)) * h.x + fabsf(dot(col1(m), axis)) * h.y +
           fabsf(dot(col2(m), axis)) * h.z;
}

// boxSupportPoint, narrowphase.cpp:27-35.
__device__ __forceinline__ f3 box_support(f3 h, const pose& t, f3 dir)
{
    const f3 l = rotate(conj(t.q), dir);
    const f3 p{l.x >= 0.0f ? h.x : -h.x, l.y >= 0.0f ? h.y : -h.y, l.z >= 0.0f ? h.z : -h.z};
    return to_world(p, t);
}

// overlapMidpoint, narrowphase.cpp:48-58.
__device__ __forceinline__ float overlap_mid(f3 ha, const pose& ta, const m3& ma, f3 hb,
                                             const pose& tb, const m3& mb, f3 axis)
{
    const float ca = dot(ta.p, axis);
    const float cb = dot(tb.p, axis);
    const float ra = proj_radius(ha, ma, axis);
    const float rb = proj_radius(hb, mb, axis);
    const float lower = max_ref(ca - ra, cb - rb);
    const float upper = min_ref(ca + ra, cb + rb);
    return (lower + upper) * 0.5f;
}

// overlapOnAxis, sat.cpp:45-62.  Returns false on separation; sets `used`
// when the axis is not degenerate (then depth/unit are valid).
__device__ __forceinline__ bool axis_overlap(f3 ha, const m3& ma, f3 hb, const m3& mb, f3 d_ab,
                                             f3 axis, f3& unit, float& depth, bool& used)
{
    const float lsq = dot(axis, axis);
    if (lsq < 1e-8f) {
        used = false;
        return true;
    }
    used = true;
    unit = scale(axis, 1.0f / sqrtf(lsq));
    const float dist = fabsf(dot(d_ab, unit));
    depth = (proj_radius(ha, ma, unit) + proj_radius(hb, mb, unit)) - dist;
    return !(depth < 0.0f);
}

// testOBBOBB, sat.cpp:66-102: A axes, B axes, then Ai x Bj (i outer); the
// strictly smaller depth wins.
__device__ bool sat(f3 ha, const pose& ta, const m3& ma, f3 hb, const pose& tb, const m3& mb,
                    f3& axis_out, float& depth_out, uint32_t& sel)
{
    sel = 0;
    depth_out = 3.402823466e+38f;
    const f3 d_ab = sub(tb.p, ta.p);
    const f3 ea[3] = {col0(ma), col1(ma), col2(ma)};
    const f3 eb[3] = {col0(mb), col1(mb), col2(mb)};
    f3 unit;
    float depth;
    bool used;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (!axis_overlap(ha, ma, hb, mb, d_ab, ea[i], unit, depth, used))
            return false;
        if (used && depth < depth_out) {
            depth_out = depth;
            axis_out = unit;
            sel = i;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (!axis_overlap(ha, ma, hb, mb, d_ab, eb[i], unit, depth, used))
            return false;
        if (used && depth < depth_out) {
            depth_out = depth;
            axis_out = unit;
            sel = 3 + i;
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            if (!axis_overlap(ha, ma, hb, mb, d_ab, cross(ea[i], eb[j]), unit, depth, used))
                return false;
            if (used && depth < depth_out) {
                depth_out = depth;
                axis_out = unit;
                sel = 6 + 3 * i + j;
            }
        }
    return true;
}

// supportFaceVertices, narrowphase.cpp:60-98 (vertex k: bit 0 -> first
// tangent axis sign, bit 1 -> second).
__device__ __forceinline__ void support_face(f3 h, const pose& t, f3 dir, f3 out[4])
{
    const f3 l = rotate(conj(t.q), dir);
    int face = 0;
    float largest = fabsf(l.x);
    if (fabsf(l.y) > largest) {
        face = 1;
        largest = fabsf(l.y);
    }
    if (fabsf(l.z) > largest)
        face = 2;
    const float comp = face == 0 ? l.x : (face == 1 ? l.y : l.z);
    const float sign = comp >= 0.0f ? 1.0f : -1.0f;
    const float ext[3] = {h.x, h.y, h.z};
    const int u = (face + 1) % 3, v = (face + 2) % 3;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float c[3] = {0.0f, 0.0f, 0.0f};
        c[face] = sign * ext[face];
        c[u] = (k & 1) ? ext[u] : -ext[u];
        c[v] = (k & 2) ? ext[v] : -ext[v];
        out[k] = to_world(f3{c[0], c[1], c[2]}, t);
    }
}

// pointInsideBox, narrowphase.cpp:100-107.
__device__ __forceinline__ bool inside_box(f3 p, f3 h, const pose& t)
{
    const f3 l = rotate(conj(t.q), sub(p, t.p));
    return fabsf(l.x) <= h.x + 1e-4f && fabsf(l.y) <= h.y + 1e-4f && fabsf(l.z) <= h.z + 1e-4f;
}

// addFaceCandidates lambda, narrowphase.cpp:281-304.
__device__ __forceinline__ void face_candidates(f3 hs, const pose& ts, f3 ho, const pose& to,
                                                f3 dir, f3 pts[4], float depths[4],
                                                uint32_t& count, BoxTape* tape, uint8_t side)
{
    const float other_surface = dot(box_support(ho, to, neg(dir)), dir);
    f3 verts[4];
    support_face(hs, ts, dir, verts);
    for (int k = 0; k < 4; ++k) {
        if (!inside_box(verts[k], ho, to))
            continue;
        const float excess = dot(verts[k], dir) - other_surface;
        const float d = max_ref(excess, 0.0f);
        const f3 cand = sub(verts[k], scale(dir, d * 0.5f));
        if (count < 4) {
            bool dup = false;
            for (uint32_t i = 0; i < count; ++i) {
                const f3 e = sub(pts[i], cand);
                if (dot(e, e) <= 1e-6f)
                    dup = true;
            }
            if (!dup) {
                if (tape) {
                    tape->side[count] = side;
                    tape->corner[count] = uint8_t(k);
                    tape->zero[count] = excess < 0.0f;
                }
                depths[count] = d;
                pts[count++] = cand;
            }
        }
    }
}

// intersectBoxBox, narrowphase.cpp:250-324.
__device__ bool box_box(f3 ha, const pose& ta, f3 hb, const pose& tb, Contact& c,
                        BoxTape* tape = nullptr)
{
    const m3 ma = to_mat(ta.q), mb = to_mat(tb.q);
    f3 n;
    float pen;
    uint32_t sel;
    if (!sat(ha, ta, ma, hb, tb, mb, n, pen, sel))
        return false;
    const f3 center_delta = sub(tb.p, ta.p);
    const bool flip = dot(center_delta, n) < 0.0f;
    if (flip)
        n = neg(n);
    if (tape) {
        tape->sel = sel;
        tape->flip = flip;
    }
    c.normal = n;
    const f3 ref = fabsf(n.x) < 0.9f ? f3{1.0f, 0.0f, 0.0f} : f3{0.0f, 1.0f, 0.0f};
    const f3 t1 = normalize(cross(n, ref));
    const f3 t2 = cross(n, t1);
    const float nc = (dot(ta.p, n) + proj_radius(ha, ma, n) + dot(tb.p, n) -
                      proj_radius(hb, mb, n)) * 0.5f;
    const f3 fallback = add(add(scale(n, nc), scale(t1, overlap_mid(ha, ta, ma, hb, tb, mb, t1))),
                            scale(t2, overlap_mid(ha, ta, ma, hb, tb, mb, t2)));
    f3 pts[4];
    float depths[4];
    uint32_t count = 0;
    face_candidates(ha, ta, hb, tb, n, pts, depths, count, tape, 0);
    face_candidates(hb, tb, ha, ta, neg(n), pts, depths, count, tape, 1);
    const bool use_fallback = count == 0;
    if (use_fallback) {
        depths[0] = pen;
        pts[0] = fallback;
        count = 1;
    }
    if (tape) {
        tape->fallback = use_fallback;
        tape->count = count;
    }
    c.count = count;
    for (uint32_t k = 0; k < count; ++k) {
        c.depth[k] = depths[k];
        set_anchors(c, k, pts[k], ta, tb);
    }
    return true;
}

}  // namespace
}  // namespace pg
