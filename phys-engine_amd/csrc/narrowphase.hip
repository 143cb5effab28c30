// Narrow phase (src/collision/narrowphase.cpp:16-374, sat.cpp:38-102,
// gjk.cpp:11-158, epa.cpp:29-130) and the world-level wrapper
// (src/world/physics_world.cpp:645-736).
//
// Pipeline, all on one stream, driven by device-side counts:
//   classify   candidate k -> shape-pair class (SS, SB, BB, SH, BH, HH) and a
//              bucket of candidate indices per class (wavefront ballot, one
//              atomic per wave and class), so each class kernel runs one code
//              path without divergence between shape types;
//   per class  one thread per candidate; the manifold goes to the candidate's
//              own slot (the reference's per-candidate result slots,
//              physics_world.cpp:727-733) and a hit flag;
//   compaction tile hit counts, one-workgroup scan, ordered gather: manifolds
//              come out in candidate order like appendResult (:692-712).
// All arithmetic replays the reference's FP32 operation order (devmath.hpp).
#include "devmath.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace pg {
namespace {

struct Bodies {
    uint32_t n;
    const float* pose;
    const uint8_t* type;
    const float* sdata;
    const uint32_t* hrange;
    const float* hpts;
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
    const uint32_t n = B.n;
    return {{B.pose[i], B.pose[n + i], B.pose[2 * n + i]},
            {B.pose[3 * n + i], B.pose[4 * n + i], B.pose[5 * n + i], B.pose[6 * n + i]}};
}
__device__ __forceinline__ f3 load_data(const Bodies& B, uint32_t i)
{
    return {B.sdata[i], B.sdata[B.n + i], B.sdata[2 * B.n + i]};
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
                    f3& axis_out, float& depth_out)
{
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
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (!axis_overlap(ha, ma, hb, mb, d_ab, eb[i], unit, depth, used))
            return false;
        if (used && depth < depth_out) {
            depth_out = depth;
            axis_out = unit;
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            if (!axis_overlap(ha, ma, hb, mb, d_ab, cross(ea[i], eb[j]), unit, depth, used))
                return false;
            if (used && depth < depth_out) {
                depth_out = depth;
                axis_out = unit;
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
                                                uint32_t& count)
{
    const float other_surface = dot(box_support(ho, to, neg(dir)), dir);
    f3 verts[4];
    support_face(hs, ts, dir, verts);
    for (int k = 0; k < 4; ++k) {
        if (!inside_box(verts[k], ho, to))
            continue;
        const float d = max_ref(dot(verts[k], dir) - other_surface, 0.0f);
        const f3 cand = sub(verts[k], scale(dir, d * 0.5f));
        if (count < 4) {
            bool dup = false;
            for (uint32_t i = 0; i < count; ++i) {
                const f3 e = sub(pts[i], cand);
                if (dot(e, e) <= 1e-6f)
                    dup = true;
            }
            if (!dup) {
                depths[count] = d;
                pts[count++] = cand;
            }
        }
    }
}

// intersectBoxBox, narrowphase.cpp:250-324.
__device__ bool box_box(f3 ha, const pose& ta, f3 hb, const pose& tb, Contact& c)
{
    const m3 ma = to_mat(ta.q), mb = to_mat(tb.q);
    f3 n;
    float pen;
    if (!sat(ha, ta, ma, hb, tb, mb, n, pen))
        return false;
    const f3 center_delta = sub(tb.p, ta.p);
    if (dot(center_delta, n) < 0.0f)
        n = neg(n);
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
    face_candidates(ha, ta, hb, tb, n, pts, depths, count);
    face_candidates(hb, tb, ha, ta, neg(n), pts, depths, count);
    if (count == 0) {
        depths[0] = pen;
        pts[0] = fallback;
        count = 1;
    }
    c.count = count;
    for (uint32_t k = 0; k < count; ++k) {
        c.depth[k] = depths[k];
        set_anchors(c, k, pts[k], ta, tb);
    }
    return true;
}

// ------------------------------------------------------------ GJK / EPA
constexpr uint32_t kEpaMaxIterations = PHYSGPU_EPA_MAX_ITERATIONS;
constexpr uint32_t kMaxVertices = 4 + kEpaMaxIterations;
constexpr uint32_t kMaxFaces = PHYSGPU_EPA_MAX_FACES;
constexpr uint32_t kMaxEdges = PHYSGPU_EPA_MAX_HORIZON;
constexpr uint32_t kGjkGrid = 1024;   // bounds the private (scratch) polytope storage

struct Shape {
    uint32_t type;
    f3 data;
    uint32_t first, count;
};

// Support maps: sphere pos + normalize(d) r; box boxSupportPoint; hull the
// first point of maximal dot with the direction in collider-local space.
__device__ f3 support(const Bodies& B, const Shape& s, const pose& t, f3 d)
{
    if (s.type == PHYSGPU_SPHERE)
        return add(t.p, scale(normalize(d), s.data.x));
    if (s.type == PHYSGPU_BOX)
        return box_support(s.data, t, d);
    const f3 l = rotate(conj(t.q), d);
    const uint32_t nh = B.nh;
    uint32_t best = s.first;
    float best_dot = dot(f3{B.hpts[best], B.hpts[nh + best], B.hpts[2 * nh + best]}, l);
    for (uint32_t k = 1; k < s.count; ++k) {
        const uint32_t p = s.first + k;
        const float v = dot(f3{B.hpts[p], B.hpts[nh + p], B.hpts[2 * nh + p]}, l);
        if (v > best_dot) {
            best_dot = v;
            best = p;
        }
    }
    return to_world(f3{B.hpts[best], B.hpts[nh + best], B.hpts[2 * nh + best]}, t);
}

struct Simplex {
    f3 p[4];
    uint32_t size;
};

// setLineDirection, gjk.cpp:21-41.
__device__ void gjk_line(Simplex& s, f3& dir)
{
    const f3 a = s.p[0], b = s.p[1];
    const f3 ao = neg(a);
    const f3 ab = sub(b, a);
    if (dot(ab, ao) > 0.0f) {
        dir = cross(cross(ab, ao), ab);
        if (dot(dir, dir) < 1e-12f)
            dir = cross(ab, f3{1.0f, 0.0f, 0.0f});
        if (dot(dir, dir) < 1e-12f)
            dir = cross(ab, f3{0.0f, 1.0f, 0.0f});
    } else {
        s.size = 1;
        dir = ao;
    }
}

// setTriangleDirection, gjk.cpp:43-74 (the line branch keeps size 3).
__device__ void gjk_triangle(Simplex& s, f3& dir)
{
    const f3 a = s.p[0], b = s.p[1], c = s.p[2];
    const f3 ao = neg(a);
    const f3 ab = sub(b, a), ac = sub(c, a);
    const f3 nrm = cross(ab, ac);
    if (dot(cross(nrm, ac), ao) > 0.0f) {
        if (dot(ac, ao) > 0.0f) {
            s.p[1] = c;
            s.size = 2;
            dir = cross(cross(ac, ao), ac);
            return;
        }
        gjk_line(s, dir);
        return;
    }
    if (dot(cross(ab, nrm), ao) > 0.0f) {
        gjk_line(s, dir);
        return;
    }
    if (dot(nrm, ao) > 0.0f) {
        dir = nrm;
    } else {
        const f3 tmp = s.p[1];
        s.p[1] = s.p[2];
        s.p[2] = tmp;
        dir = neg(nrm);
    }
}

// containsOrigin, gjk.cpp:76-124; in the third face case the reference
// stores D into points[1] and then reads points[1] for points[2] (both D).
__device__ bool gjk_contains(Simplex& s, f3& dir)
{
    if (s.size == 1) {
        dir = neg(s.p[0]);
        return false;
    }
    if (s.size == 2) {
        gjk_line(s, dir);
        return false;
    }
    if (s.size == 3) {
        gjk_triangle(s, dir);
        return false;
    }
    const f3 a = s.p[0];
    const f3 ao = neg(a);
    auto outside = [&](f3 fb, f3 fc, f3 opp) {
        f3 nrm = cross(sub(fb, a), sub(fc, a));
        if (dot(nrm, sub(opp, a)) > 0.0f)
            nrm = neg(nrm);
        return dot(nrm, ao) > 0.0f;
    };
    if (outside(s.p[1], s.p[2], s.p[3])) {
        s.size = 3;
        gjk_triangle(s, dir);
        return false;
    }
    if (outside(s.p[2], s.p[3], s.p[1])) {
        s.p[1] = s.p[2];
        s.p[2] = s.p[3];
        s.size = 3;
        gjk_triangle(s, dir);
        return false;
    }
    if (outside(s.p[3], s.p[1], s.p[2])) {
        s.p[1] = s.p[3];
        s.p[2] = s.p[1];
        s.size = 3;
        gjk_triangle(s, dir);
        return false;
    }
    return true;
}

// intersectGjk, gjk.cpp:128-158.
__device__ bool gjk(const Bodies& B, const Shape& sa, const pose& ta, const Shape& sb,
                    const pose& tb, uint32_t iterations, Simplex& s)
{
    s.size = 0;
    f3 dir{1.0f, 0.0f, 0.0f};
    for (uint32_t it = 0; it < iterations; ++it) {
        if (dot(dir, dir) < 1e-12f)
            break;
        const f3 p = sub(support(B, sa, ta, dir), support(B, sb, tb, neg(dir)));
        if (dot(p, dir) <= 0.0f)
            break;
        for (uint32_t k = min(s.size, 3u); k > 0; --k)
            s.p[k] = s.p[k - 1];
        s.p[0] = p;
        s.size = min(s.size + 1, 4u);
        if (gjk_contains(s, dir))
            return true;
    }
    return false;
}

struct Face {
    uint8_t a, b, c, pad;
    f3 n;
    float d;
};

// makeFace, epa.cpp:29-47.
__device__ __forceinline__ Face make_face(const f3* v, uint32_t a, uint32_t b, uint32_t c)
{
    f3 n = cross(sub(v[b], v[a]), sub(v[c], v[a]));
    const float len = length(n);
    if (len > 1e-8f)
        n = scale(n, 1.0f / len);
    float d = dot(n, v[a]);
    if (d < 0.0f) {
        const uint32_t t = b;
        b = c;
        c = t;
        n = neg(n);
        d = -d;
    }
    return {uint8_t(a), uint8_t(b), uint8_t(c), 0, n, d};
}

// computeEpa, epa.cpp:64-130, with bounded polytope storage.
// Returns 1 converged, 0 failed (reference failure), -1 storage exceeded.
__device__ int epa(const Bodies& B, const Shape& sa, const pose& ta, const Shape& sb,
                   const pose& tb, const Simplex& s, uint32_t iterations, float tol, f3& normal,
                   float& depth)
{
    normal = {0.0f, 0.0f, 0.0f};
    depth = 0.0f;
    if (s.size < 4 || iterations == 0)
        return 0;
    f3 v[kMaxVertices];
    Face faces[kMaxFaces];
    uint8_t ea[kMaxEdges], eb[kMaxEdges];
    uint32_t nv = 4, nf = 4;
    for (int k = 0; k < 4; ++k)
        v[k] = s.p[k];
    faces[0] = make_face(v, 0, 1, 2);
    faces[1] = make_face(v, 0, 3, 1);
    faces[2] = make_face(v, 0, 2, 3);
    faces[3] = make_face(v, 1, 3, 2);
    for (uint32_t it = 0; it < iterations; ++it) {
        if (nf == 0)
            return 0;
        uint32_t best = 0;
        for (uint32_t f = 1; f < nf; ++f)
            if (faces[f].d < faces[best].d)
                best = f;
        const Face cf = faces[best];
        const f3 sup = sub(support(B, sa, ta, cf.n), support(B, sb, tb, neg(cf.n)));
        if (dot(sup, cf.n) - cf.d <= tol) {
            normal = cf.n;
            depth = max_ref(cf.d, 0.0f);
            return 1;
        }
        if (nv >= kMaxVertices)
            return -1;
        const uint32_t si = nv;
        v[nv++] = sup;
        uint32_t ne = 0;
        uint32_t keep = 0;
        bool overflow = false;
        auto add_edge = [&](uint32_t a, uint32_t b) {
            for (uint32_t e = 0; e < ne; ++e)
                if (ea[e] == b && eb[e] == a) {
                    for (uint32_t k = e + 1; k < ne; ++k) {
                        ea[k - 1] = ea[k];
                        eb[k - 1] = eb[k];
                    }
                    --ne;
                    return;
                }
            if (ne < kMaxEdges) {
                ea[ne] = uint8_t(a);
                eb[ne] = uint8_t(b);
                ++ne;
            } else {
                overflow = true;  // the unbounded horizon would grow past kMaxEdges
            }
        };
        for (uint32_t f = 0; f < nf; ++f) {
            const Face fc = faces[f];
            if (dot(fc.n, sub(sup, v[fc.a])) > tol) {
                add_edge(fc.a, fc.b);
                add_edge(fc.b, fc.c);
                add_edge(fc.c, fc.a);
            } else {
                faces[keep++] = fc;
            }
        }
        if (overflow)
            return -1;
        if (ne == 0)
            return 0;
        if (keep + ne > kMaxFaces)
            return -1;
        nf = keep;
        for (uint32_t e = 0; e < ne; ++e)
            faces[nf++] = make_face(v, ea[e], eb[e], si);
    }
    return 0;
}

// ------------------------------------------------------------ kernels
__device__ __forceinline__ void wave_append_u32(bool pred, uint32_t value, uint32_t* out,
                                                uint32_t* counter)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0)
        return;
    const uint32_t lane = lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    if (pred)
        out[base + __popcll(mask & lanes_below(lane))] = value;
}

__device__ __forceinline__ void wave_count(bool pred, uint32_t* counter)
{
    const uint64_t mask = __ballot(pred);
    if (mask && lane_id() == (uint32_t)__ffsll((unsigned long long)mask) - 1u)
        atomicAdd(counter, (uint32_t)__popcll(mask));
}

__global__ __launch_bounds__(kBlock) void k_classify(Bodies B, World W,
                                                     const uint64_t* __restrict__ pairs,
                                                     Counters* c, uint32_t* __restrict__ buckets,
                                                     uint64_t cap, uint32_t* __restrict__ hit)
{
    const uint32_t count = c->sort_count;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (count + stride - 1) / stride;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t k = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        uint32_t cls = kClasses;
        if (k < count) {
            const uint64_t pr = pairs[k];
            const uint32_t a = uint32_t(pr >> 32), b = uint32_t(pr);
            hit[k] = 0;
            // Same-body pairs are removed before the narrow phase
            // (physics_world.cpp:630-633).
            if (!(W.body && W.body[a] == W.body[b]))
                cls = pair_class(B.type[a], B.type[b]);
        }
#pragma unroll
        for (uint32_t q = 0; q < kClasses; ++q)
            wave_append_u32(cls == q, k, buckets + q * cap, &c->class_count[q]);
    }
}

// Per-candidate record: collider-local anchors, or body-local ones plus the
// combined materials at world level (physics_world.cpp:679-690).
__device__ void store_contact(const Contact& ct, uint32_t k, uint32_t a, uint32_t b,
                              uint32_t cls, const World& W, uint32_t n,
                              physgpu_manifold* __restrict__ slots)
{
    uint32_t w[40];
    w[0] = a;
    w[1] = b;
    w[2] = __float_as_uint(ct.normal.x);
    w[3] = __float_as_uint(ct.normal.y);
    w[4] = __float_as_uint(ct.normal.z);
    w[5] = ct.count;
    pose la{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 1.0f}}, lb = la;
    const bool local = W.local_pose != nullptr;
    if (local) {
        const float* L = W.local_pose;
        la = {{L[a], L[n + a], L[2 * n + a]}, {L[3 * n + a], L[4 * n + a], L[5 * n + a], L[6 * n + a]}};
        lb = {{L[b], L[n + b], L[2 * n + b]}, {L[3 * n + b], L[4 * n + b], L[5 * n + b], L[6 * n + b]}};
    }
#pragma unroll
    for (uint32_t p = 0; p < 4; ++p) {
        f3 pa{0.0f, 0.0f, 0.0f}, pb{0.0f, 0.0f, 0.0f};
        float d = 0.0f;
        if (p < ct.count) {
            pa = ct.a[p];
            pb = ct.b[p];
            d = ct.depth[p];
            if (local) {
                pa = add(la.p, rotate(la.q, pa));
                pb = add(lb.p, rotate(lb.q, pb));
            }
        }
        w[6 + 3 * p] = __float_as_uint(pa.x);
        w[7 + 3 * p] = __float_as_uint(pa.y);
        w[8 + 3 * p] = __float_as_uint(pa.z);
        w[18 + 3 * p] = __float_as_uint(pb.x);
        w[19 + 3 * p] = __float_as_uint(pb.y);
        w[20 + 3 * p] = __float_as_uint(pb.z);
        w[30 + p] = __float_as_uint(d);
    }
    float fr = 0.0f, rs = 0.0f;
    if (W.friction)
        fr = sqrtf(max_ref(W.friction[a], 0.0f) * max_ref(W.friction[b], 0.0f));
    if (W.restitution)
        rs = max_ref(W.restitution[a], W.restitution[b]);
    w[34] = __float_as_uint(fr);
    w[35] = __float_as_uint(rs);
    w[36] = k;
    w[37] = cls;
    w[38] = 0;
    w[39] = 0;
    uint4* dst = reinterpret_cast<uint4*>(slots + k);
#pragma unroll
    for (int q = 0; q < 10; ++q)
        dst[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

template <int kClass>
__global__ __launch_bounds__(kBlock) void k_class(Bodies B, World W,
                                                  const uint64_t* __restrict__ pairs, Counters* c,
                                                  const uint32_t* __restrict__ buckets,
                                                  uint64_t cap, physgpu_narrowphase_options opt,
                                                  physgpu_manifold* __restrict__ slots,
                                                  uint32_t* __restrict__ hit)
{
    // Classes >= 3 (hull pairs) share one launch: buckets 3, 4, 5 in turn.
    uint32_t total = c->class_count[kClass];
    if (kClass == 3)
        total += c->class_count[4] + c->class_count[5];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (total + stride - 1) / stride;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t t = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool h = false;
        bool failed = false, overflowed = false;
        uint32_t cls = kClass;
        if (t < total) {
            uint32_t q = t;
            if (kClass == 3) {
                const uint32_t c3 = c->class_count[3], c4 = c->class_count[4];
                if (q >= c3 + c4) {
                    cls = 5;
                    q -= c3 + c4;
                } else if (q >= c3) {
                    cls = 4;
                    q -= c3;
                }
            }
            const uint32_t k = buckets[cls * cap + q];
            const uint64_t pr = pairs[k];
            const uint32_t a = uint32_t(pr >> 32), b = uint32_t(pr);
            const pose ta = load_pose(B, a), tb = load_pose(B, b);
            Contact ct;
            if (kClass == PHYSGPU_PAIR_SPHERE_SPHERE) {
                h = sphere_sphere(B.sdata[a], ta, B.sdata[b], tb, ct);
            } else if (kClass == PHYSGPU_PAIR_SPHERE_BOX) {
                if (B.type[a] == PHYSGPU_SPHERE) {
                    h = sphere_box(B.sdata[a], ta, load_data(B, b), tb, ct);
                } else {
                    // Box-sphere: swap, then negate the normal and swap anchors (:353-365).
                    h = sphere_box(B.sdata[b], tb, load_data(B, a), ta, ct);
                    if (h) {
                        ct.normal = neg(ct.normal);
                        const f3 tmp = ct.a[0];
                        ct.a[0] = ct.b[0];
                        ct.b[0] = tmp;
                    }
                }
            } else if (kClass == PHYSGPU_PAIR_BOX_BOX) {
                h = box_box(load_data(B, a), ta, load_data(B, b), tb, ct);
            } else {
                const uint32_t n = B.n;
                Shape sa{B.type[a], load_data(B, a), 0, 0}, sb{B.type[b], load_data(B, b), 0, 0};
                if (sa.type == PHYSGPU_HULL) {
                    sa.first = B.hrange[a];
                    sa.count = B.hrange[n + a];
                }
                if (sb.type == PHYSGPU_HULL) {
                    sb.first = B.hrange[b];
                    sb.count = B.hrange[n + b];
                }
                Simplex s;
                if (gjk(B, sa, ta, sb, tb, opt.gjk_max_iterations, s)) {
                    f3 nrm;
                    float depth;
                    const int st = epa(B, sa, ta, sb, tb, s, opt.epa_max_iterations,
                                       opt.epa_tolerance, nrm, depth);
                    if (st == 1) {
                        h = true;
                        ct.normal = nrm;
                        ct.count = 1;
                        ct.depth[0] = depth;
                        const f3 w = scale(add(support(B, sa, ta, nrm), support(B, sb, tb, neg(nrm))),
                                           0.5f);
                        set_anchors(ct, 0, w, ta, tb);
                    } else {
                        failed = st == 0;
                        overflowed = st < 0;
                    }
                }
            }
            if (h) {
                store_contact(ct, k, a, b, cls, W, B.n, slots);
                hit[k] = 1;
            }
        }
        if (kClass == 3) {
            for (uint32_t q = 3; q < kClasses; ++q)
                wave_count(h && cls == q, &c->class_contacts[q]);
            wave_count(failed, &c->gjk_failures);
            wave_count(overflowed, &c->epa_overflow);
        } else {
            wave_count(h, &c->class_contacts[kClass]);
        }
    }
}

// Compaction: hits per tile of kScanTile candidates.
__global__ __launch_bounds__(kBlock) void k_tile_hits(const uint32_t* __restrict__ hit,
                                                      const Counters* c,
                                                      uint32_t* __restrict__ tile_sums)
{
    __shared__ uint32_t part[kBlock / 64];
    const uint32_t count = c->sort_count;
    const uint32_t tiles = (count + kScanTile - 1) / kScanTile;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        uint32_t s = 0;
        for (uint32_t i = threadIdx.x; i < kScanTile; i += blockDim.x) {
            const uint32_t k = tile * kScanTile + i;
            if (k < count)
                s += hit[k];
        }
        for (uint32_t off = 32; off > 0; off >>= 1)
            s += __shfl_xor(s, off);
        __syncthreads();
        if (lane_id() == 0)
            part[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < kBlock / 64; ++w)
                t += part[w];
            tile_sums[tile] = t;
        }
    }
}

__global__ __launch_bounds__(1024) void k_tile_scan(uint32_t* __restrict__ tile_sums, Counters* c)
{
    __shared__ uint32_t sums[1024];
    const uint32_t count = c->sort_count;
    const uint32_t tiles = (count + kScanTile - 1) / kScanTile;
    const uint32_t per = (tiles + 1023) / 1024;
    const uint32_t begin = min(threadIdx.x * per, tiles), end = min(begin + per, tiles);
    uint32_t s = 0;
    for (uint32_t i = begin; i < end; ++i)
        s += tile_sums[i];
    sums[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = threadIdx.x >= off ? sums[threadIdx.x - off] : 0u;
        __syncthreads();
        sums[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? sums[threadIdx.x - 1] : 0u;
    for (uint32_t i = begin; i < end; ++i) {
        const uint32_t v = tile_sums[i];
        tile_sums[i] = run;
        run += v;
    }
    if (threadIdx.x == 1023)
        c->manifold_count = sums[1023];
}

// Ordered gather: thread t of a tile owns 8 consecutive candidates.
__global__ __launch_bounds__(kBlock) void k_compact(const uint32_t* __restrict__ hit,
                                                    const Counters* c,
                                                    const uint32_t* __restrict__ tile_sums,
                                                    const physgpu_manifold* __restrict__ slots,
                                                    physgpu_manifold* __restrict__ out)
{
    __shared__ uint32_t wave_tot[kBlock / 64];
    constexpr uint32_t per = kScanTile / kBlock;
    const uint32_t count = c->sort_count;
    const uint32_t tiles = (count + kScanTile - 1) / kScanTile;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t k0 = tile * kScanTile + threadIdx.x * per;
        uint32_t flags = 0, mine = 0;
        for (uint32_t i = 0; i < per; ++i) {
            const uint32_t k = k0 + i;
            const uint32_t f = (k < count) ? hit[k] : 0u;
            flags |= f << i;
            mine += f;
        }
        uint32_t incl = mine;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off);
            if (lane >= off)
                incl += t;
        }
        __syncthreads();
        if (lane == 63)
            wave_tot[wave] = incl;
        __syncthreads();
        uint32_t base = tile_sums[tile];
        for (uint32_t w = 0; w < wave; ++w)
            base += wave_tot[w];
        uint32_t pos = base + incl - mine;
        for (uint32_t i = 0; i < per; ++i) {
            if (flags & (1u << i)) {
                const uint4* src = reinterpret_cast<const uint4*>(slots + k0 + i);
                uint4* dst = reinterpret_cast<uint4*>(out + pos);
#pragma unroll
                for (int q = 0; q < 10; ++q)
                    dst[q] = src[q];
                ++pos;
            }
        }
    }
}

__global__ void k_np_begin(Counters* c)
{
    if (threadIdx.x < kClasses) {
        c->class_count[threadIdx.x] = 0;
        c->class_contacts[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) {
        c->gjk_failures = 0;
        c->epa_overflow = 0;
        c->manifold_count = 0;
    }
}

}  // namespace

void launch_np_begin(physgpu_context* ctx)
{
    hipLaunchKernelGGL(k_np_begin, dim3(1), dim3(64), 0, ctx->stream, ctx->counters);
}

void launch_narrowphase(physgpu_context* ctx, bool world)
{
    Bodies B{ctx->n, ctx->pose, ctx->shape_type, ctx->shape_data, ctx->hull_range,
             ctx->hull_points, ctx->n_hull_points};
    World W{nullptr, nullptr, nullptr, nullptr};
    if (world && ctx->has_world)
        W = World{ctx->world_has_body ? ctx->body : nullptr,
                  ctx->world_has_local ? ctx->local_pose : nullptr,
                  ctx->world_has_friction ? ctx->friction : nullptr,
                  ctx->world_has_restitution ? ctx->restitution : nullptr};
    physgpu_narrowphase_options opt = ctx->np_options;
    const uint32_t grid = kStrideGrid;
    Timer t;
    time_begin(ctx, "np_classify", t);
    hipLaunchKernelGGL(k_classify, dim3(grid), dim3(kBlock), 0, ctx->stream, B, W, ctx->pairs,
                       ctx->counters, ctx->buckets, ctx->cap_p, ctx->hit);
    time_end(ctx, t);
    time_begin(ctx, "np_sphere_sphere", t);
    hipLaunchKernelGGL(k_class<0>, dim3(grid), dim3(kBlock), 0, ctx->stream, B, W, ctx->pairs,
                       ctx->counters, ctx->buckets, ctx->cap_p, opt, ctx->slots, ctx->hit);
    time_end(ctx, t);
    time_begin(ctx, "np_sphere_box", t);
    hipLaunchKernelGGL(k_class<1>, dim3(grid), dim3(kBlock), 0, ctx->stream, B, W, ctx->pairs,
                       ctx->counters, ctx->buckets, ctx->cap_p, opt, ctx->slots, ctx->hit);
    time_end(ctx, t);
    time_begin(ctx, "np_box_box", t);
    hipLaunchKernelGGL(k_class<2>, dim3(grid), dim3(kBlock), 0, ctx->stream, B, W, ctx->pairs,
                       ctx->counters, ctx->buckets, ctx->cap_p, opt, ctx->slots, ctx->hit);
    time_end(ctx, t);
    if (ctx->has_hulls) {
        time_begin(ctx, "np_gjk_epa", t);
        hipLaunchKernelGGL(k_class<3>, dim3(kGjkGrid), dim3(kBlock), 0, ctx->stream, B, W,
                           ctx->pairs, ctx->counters, ctx->buckets, ctx->cap_p, opt, ctx->slots,
                           ctx->hit);
        time_end(ctx, t);
    }
    time_begin(ctx, "np_compact", t);
    hipLaunchKernelGGL(k_tile_hits, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->hit,
                       ctx->counters, ctx->tile_sums);
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, ctx->stream, ctx->tile_sums,
                       ctx->counters);
    hipLaunchKernelGGL(k_compact, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->hit,
                       ctx->counters, ctx->tile_sums, ctx->slots, ctx->manifolds);
    time_end(ctx, t);
}

}  // namespace pg
