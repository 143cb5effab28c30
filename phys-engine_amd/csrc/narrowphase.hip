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
#include <stddef.h>

#include <algorithm>

#include "contact.hpp"
#include "devmath.hpp"
#include "fold.hpp"
#include "internal.hpp"
#include "vjp.hpp"
#include "wave.hpp"

namespace pg {
namespace {

// ------------------------------------------------------------ GJK / EPA
constexpr uint32_t kEpaMaxIterations = PHYSGPU_EPA_MAX_ITERATIONS;
constexpr uint32_t kMaxVertices = 4 + kEpaMaxIterations;
constexpr uint32_t kMaxFaces = PHYSGPU_EPA_MAX_FACES;
constexpr uint32_t kMaxEdges = PHYSGPU_EPA_MAX_HORIZON;
// Work items finished by GJK + EPA (kSlotEpaItemsDone), counted on
// kEpaDoneSlots addresses one cache line apart.
constexpr uint32_t kEpaDoneSlots = 64;
constexpr uint32_t kEpaDoneStride = 16;
static_assert(kEpaDoneSlots * kEpaDoneStride <= kEpaPartials, "done counts lie in the zeroed slot");
static_assert(kSlotGjkFailures == kEpaPartialFirst + 3 && kSlotEpaOverflows == kEpaPartialFirst + 4 &&
                  kSlotEpaItemsDone == kEpaPartialFirst + 5,
              "EPA partial slots are contiguous");

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

// computeEpa (epa.cpp:64-130) run by one wavefront per pair: the polytope
// lives in LDS, faces are spread over the 64 lanes (closest face = wave
// arg-min with the reference's first-minimum rule, visibility = one ballot
// per 64 faces, order-preserving in-place compaction of the kept faces), the
// horizon is rebuilt from the visible faces' edges with the reference's
// erase-the-reversed-edge rule (see horizon_wave), and the new faces are made
// in parallel, in horizon order.  Every float is computed by the same
// expression as the reference, so results are bit-identical; only the
// storage is bounded.
constexpr uint32_t kEpaWavesPerBlock = kBlock / 64;
constexpr uint32_t kLongEpa = 12;
constexpr uint32_t kVeryLongEpa = 40;
constexpr uint32_t kFaceChunks = (kMaxFaces + 63) / 64;

constexpr uint32_t kEdgeKeys = kMaxVertices * (kMaxVertices - 1) / 2;  // unordered vertex pairs
constexpr uint32_t kEdgeKeyBits = 12;
static_assert(kEdgeKeys <= (1u << kEdgeKeyBits), "edge keys must fit kEdgeKeyBits");
constexpr uint32_t kMaxVisibleEdges = 3 * kMaxFaces;
static_assert(kMaxFaces <= 256 && kMaxVertices <= 256, "face / vertex indices are bytes");

struct EpaShared {
    float vx[kMaxVertices], vy[kMaxVertices], vz[kMaxVertices];
    uint32_t fabc[kMaxFaces];  // a | b << 8 | c << 16
    float fnx[kMaxFaces], fny[kMaxFaces], fnz[kMaxFaces], fd[kMaxFaces];
    uint8_t ea[kMaxEdges], eb[kMaxEdges];
    uint8_t vis_list[kMaxFaces];                  // visible faces, in index order
    uint32_t edge_bits[kMaxVisibleEdges / 32];    // per-edge flags of the horizon passes
};

__device__ __forceinline__ uint32_t edge_key(uint32_t a, uint32_t b)
{
    const uint32_t u = a < b ? a : b, v = a < b ? b : a;
    return v * (v - 1) / 2 + u;
}

template <class Shared>
__device__ __forceinline__ f3 lds_vertex(const Shared& S, uint32_t i)
{
    return {S.vx[i], S.vy[i], S.vz[i]};
}

// makeFace, epa.cpp:29-47, over the LDS vertices; writes face slot f and
// returns its distance.
template <class Shared>
__device__ __forceinline__ float make_face_lds(Shared& S, uint32_t f, uint32_t a, uint32_t b,
                                               uint32_t c)
{
    const f3 va = lds_vertex(S, a);
    f3 n = cross(sub(lds_vertex(S, b), va), sub(lds_vertex(S, c), va));
    const float len = length(n);
    if (len > 1e-8f)
        n = scale(n, 1.0f / len);
    float d = dot(n, va);
    if (d < 0.0f) {
        const uint32_t t = b;
        b = c;
        c = t;
        n = neg(n);
        d = -d;
    }
    S.fabc[f] = a | (b << 8) | (c << 16);
    S.fnx[f] = n.x;
    S.fny[f] = n.y;
    S.fnz[f] = n.z;
    S.fd[f] = d;
    return d;
}

// A hull's points (at most 64) or a box's 8 corners held one per lane for the
// whole EPA, in collider space (p) and in world space (w = to_world(p), the
// same function of the same inputs the support map would evaluate, so the
// same bits), so support evaluations read no memory and rotate only the
// direction.  Box corner c has the +extent on the axes of its set bits
// (bit 0 x, bit 1 y, bit 2 z).
struct WaveHull {
    f3 p, w;
    bool cached;  // hull of at most 64 points
    bool box;
};

__device__ __forceinline__ WaveHull wave_hull(const Bodies& B, const Shape& s, const pose& t)
{
    WaveHull h{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, s.type == PHYSGPU_HULL && s.count <= 64,
               s.type == PHYSGPU_BOX};
    const uint32_t lane = lane_id();
    if (h.cached && lane < s.count) {
        const uint32_t p = s.first + lane, nh = B.nh;
        h.p = {B.hpts[p], B.hpts[nh + p], B.hpts[2 * nh + p]};
        h.w = to_world(h.p, t);
    } else if (h.box && lane < 8) {
        h.p = {(lane & 1u) ? s.data.x : -s.data.x, (lane & 2u) ? s.data.y : -s.data.y,
               (lane & 4u) ? s.data.z : -s.data.z};
        h.w = to_world(h.p, t);
    }
    return h;
}

__device__ __forceinline__ f3 lane_point(const f3& v, uint32_t src)
{
    return {lane_value(v.x, src), lane_value(v.y, src), lane_value(v.z, src)};
}

// Support map evaluated by the whole wave (same result in every lane).  A
// hull's points are spread over the lanes; the arg-max keeps the lowest
// index among equal dot products, as the sequential strict-> scan does.
// `local` (optional) receives the collider-space point (the box corner).
__device__ f3 support_wave(const Bodies& B, const Shape& s, const pose& t, f3 d,
                           const WaveHull& h, f3* local = nullptr)
{
    if (s.type != PHYSGPU_HULL) {
        if (h.box) {
            // boxSupportPoint's corner (narrowphase.cpp:27-35).
            const f3 l = rotate(conj(t.q), d);
            const uint32_t c = (l.x >= 0.0f ? 1u : 0u) | (l.y >= 0.0f ? 2u : 0u) |
                               (l.z >= 0.0f ? 4u : 0u);
            if (local)
                *local = lane_point(h.p, c);
            return lane_point(h.w, c);
        }
        return support(B, s, t, d);
    }
    const f3 l = rotate(conj(t.q), d);
    const uint32_t nh = B.nh, lane = lane_id();
    float bv = 0.0f;
    uint32_t bi = kNoIndex;
    if (h.cached && s.count <= 16) {
        // All points in the first 16-lane row.
        if (lane < s.count) {
            bv = dot(h.p, l);
            bi = lane;
        }
        if (!row_arg_best_lanes<true>(bv, bi, 1, 0)) {
            row_arg_best<true>(bv, bi);
            bi = lane_value(bi, 0);
        }
    } else {
        for (uint32_t base = 0; base < s.count; base += 64) {
            const uint32_t k = base + lane;
            if (k < s.count) {
                const uint32_t p = s.first + k;
                const f3 pt = h.cached ? h.p : f3{B.hpts[p], B.hpts[nh + p], B.hpts[2 * nh + p]};
                const float v = dot(pt, l);
                if (bi == kNoIndex || v > bv) {
                    bv = v;
                    bi = k;
                }
            }
        }
        if (!wave_arg_best_lanes<true>(bv, bi, (s.count + 63) / 64)) {
            wave_arg_best<true>(bv, bi);
            bi = uniform(bi);
        }
    }
    if (h.cached) {
        if (local)
            *local = lane_point(h.p, bi);
        return lane_point(h.w, bi);
    }
    const uint32_t p = s.first + bi;
    const f3 pt{B.hpts[p], B.hpts[nh + p], B.hpts[2 * nh + p]};
    if (local)
        *local = pt;
    return to_world(pt, t);
}

// Edge i of the visible faces' edge list: face vis_list[i / 3], edge i % 3
// of (a,b), (b,c), (c,a) — the order epa.cpp:106-113 feeds addHorizonEdge.
struct EdgeLane {
    uint32_t a, b, key;
    int delta;       // +1 for a < b, -1 otherwise
    bool valid;
    uint64_t same;   // lanes of this chunk holding the same vertex pair
};

__device__ __forceinline__ EdgeLane edge_fields(const EpaShared& S, uint32_t i, uint32_t edges)
{
    EdgeLane e{0, 0, 0, 0, i < edges, 0};
    if (e.valid) {
        const uint32_t r = i / 3, k = i - 3 * r;
        const uint32_t abc = S.fabc[S.vis_list[r]];
        const uint32_t k1 = k == 2 ? 0u : k + 1;
        e.a = (abc >> (8 * k)) & 0xFFu;
        e.b = (abc >> (8 * k1)) & 0xFFu;
        e.key = edge_key(e.a, e.b);
        e.delta = e.a < e.b ? 1 : -1;
    }
    return e;
}

__device__ __forceinline__ EdgeLane edge_lane(const EpaShared& S, uint32_t i, uint32_t edges)
{
    EdgeLane e = edge_fields(S, i, edges);
    e.same = match_lanes<kEdgeKeyBits>(e.key, __ballot(e.valid));
    return e;
}

// Horizons of at most this many edges (five visible faces, the common case)
// match keys by reading each edge's key from its lane instead of one ballot
// per key bit.
constexpr uint32_t kSmallHorizon = 16;

__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }

// Horizon (addHorizonEdge, epa.cpp:49-60) of `edges` visible-face edges.
// For one unordered vertex pair the reference's list only ever holds copies
// of one direction; with s = running sum of +1 (a < b) / -1 (a > b) over that
// pair's edges it holds |s| copies of direction sign(s).  An edge is appended
// exactly when it moves s away from zero (else it erases the oldest reversed
// copy), so the list size, and the point where it would pass kMaxEdges,
// follow from s alone; the survivors are the last |s_final| appended copies,
// kept in arrival order.  Edges are taken 64 per step in arrival order; lanes
// with the same pair find each other by ballots over the key bits, so sums
// and counts inside a step are popcounts.  With more than one step (over 21
// visible faces, rare), each lane also matches its key against every other
// step's keys: earlier steps give its running sum, all steps its final sum,
// later steps the appended copies that follow it.  Writes the horizon to
// S.ea/S.eb; returns its size, or -1 when the list would pass kMaxEdges.
__device__ __forceinline__ int horizon_wave(EpaShared& S, uint32_t edges)
{
    const uint32_t lane = lane_id();
    const uint64_t below = lanes_below(lane);
    const uint64_t upto = below | (1ull << lane);
    if (edges <= 64) {
        // One step: edge i in lane i; every sum is a popcount.
        EdgeLane e = edge_fields(S, lane, edges);
        if (edges <= kSmallHorizon) {
            uint64_t same = 0;
            for (uint32_t j = 0; j < edges; ++j)
                same |= lane_value(e.key, j) == e.key ? (1ull << j) : 0ull;
            e.same = same;
        } else {
            e.same = match_lanes<kEdgeKeyBits>(e.key, __ballot(e.valid));
        }
        const uint64_t up = __ballot(e.valid && e.delta > 0);
        const int sum = 2 * __popcll(e.same & up & upto) - __popcll(e.same & upto);
        const bool append = e.valid && iabs(sum) > iabs(sum - e.delta);
        const uint64_t A = __ballot(append), V = __ballot(e.valid);
        const int size_before = 2 * __popcll(A & below) - __popcll(V & below);
        if (__ballot(append && size_before >= int(kMaxEdges)))
            return -1;
        const int total = 2 * __popcll(e.same & up) - __popcll(e.same);
        const bool survive = append && __popcll(e.same & A & ~upto) < iabs(total);
        const uint64_t K = __ballot(survive);
        if (survive) {
            const uint32_t pos = (uint32_t)__popcll(K & below);
            S.ea[pos] = uint8_t(e.a);
            S.eb[pos] = uint8_t(e.b);
        }
        wave_sync();
        return __popcll(K);
    }
    // Several steps.  A lane's sums over the edges of another step come from
    // matching its key against that step's keys, one ballot per key bit.
    const uint32_t steps = (edges + 63) / 64;
    int list_size = 0;
    for (uint32_t c = 0; c < steps; ++c) {
        const EdgeLane e = edge_lane(S, c * 64 + lane, edges);
        int carry = 0;  // running sum of this lane's pair over the earlier steps
        for (uint32_t z = 0; z < c; ++z) {
            const EdgeLane f = edge_fields(S, z * 64 + lane, edges);
            const uint64_t m = match_other<kEdgeKeyBits>(e.key, f.key, __ballot(f.valid));
            const uint64_t fup = __ballot(f.valid && f.delta > 0);
            carry += 2 * __popcll(m & fup) - __popcll(m);
        }
        const uint64_t up = __ballot(e.valid && e.delta > 0);
        const int sum = carry + 2 * __popcll(e.same & up & upto) - __popcll(e.same & upto);
        const bool append = e.valid && iabs(sum) > iabs(sum - e.delta);
        const uint64_t A = __ballot(append), V = __ballot(e.valid);
        const int size_before = list_size + 2 * __popcll(A & below) - __popcll(V & below);
        if (__ballot(append && size_before >= int(kMaxEdges)))
            return -1;
        if (lane == 0) {
            S.edge_bits[2 * c] = uint32_t(A);
            S.edge_bits[2 * c + 1] = uint32_t(A >> 32);
        }
        list_size += 2 * __popcll(A) - __popcll(V);
    }
    wave_sync();
    // An appended edge survives while fewer than |final sum of its pair|
    // appended copies of the pair follow it.  Step c's flags become its
    // survivor flags; later steps still read the append flags of steps > c.
    for (uint32_t c = 0; c < steps; ++c) {
        const EdgeLane e = edge_lane(S, c * 64 + lane, edges);
        const uint64_t A =
            uint64_t(S.edge_bits[2 * c]) | (uint64_t(S.edge_bits[2 * c + 1]) << 32);
        const uint64_t up = __ballot(e.valid && e.delta > 0);
        int total = 2 * __popcll(e.same & up) - __popcll(e.same);
        int after = __popcll(e.same & A & ~upto);
        for (uint32_t z = 0; z < steps; ++z) {
            if (z == c)
                continue;
            const EdgeLane f = edge_fields(S, z * 64 + lane, edges);
            const uint64_t m = match_other<kEdgeKeyBits>(e.key, f.key, __ballot(f.valid));
            const uint64_t fup = __ballot(f.valid && f.delta > 0);
            total += 2 * __popcll(m & fup) - __popcll(m);
            if (z > c) {
                const uint64_t Az =
                    uint64_t(S.edge_bits[2 * z]) | (uint64_t(S.edge_bits[2 * z + 1]) << 32);
                after += __popcll(m & Az);
            }
        }
        const bool survive = ((A >> lane) & 1ull) && after < iabs(total);
        const uint64_t K = __ballot(survive);
        wave_sync();
        if (lane == 0) {
            S.edge_bits[2 * c] = uint32_t(K);
            S.edge_bits[2 * c + 1] = uint32_t(K >> 32);
        }
        wave_sync();
    }
    // Survivors in arrival order.
    uint32_t ne = 0;
    for (uint32_t c = 0; c < steps; ++c) {
        const uint32_t i = c * 64 + lane;
        const uint64_t K =
            uint64_t(S.edge_bits[2 * c]) | (uint64_t(S.edge_bits[2 * c + 1]) << 32);
        if ((K >> lane) & 1ull) {
            const uint32_t r = i / 3, k = i - 3 * r;
            const uint32_t abc = S.fabc[S.vis_list[r]];
            const uint32_t k1 = k == 2 ? 0u : k + 1;
            const uint32_t pos = ne + (uint32_t)__popcll(K & below);
            S.ea[pos] = uint8_t((abc >> (8 * k)) & 0xFFu);
            S.eb[pos] = uint8_t((abc >> (8 * k1)) & 0xFFu);
        }
        ne += (uint32_t)__popcll(K);
    }
    wave_sync();
    return int(ne);
}

// Returns 1 converged, 0 failed (the reference's failure), -1 storage bound
// exceeded.  `normal`/`depth` are uniform across the wave.
// On -1 from a full horizon or face list, *stop (optional) receives the
// iteration and the face count before it (it | nf << 16): S then still holds
// that iteration's faces, and its vertices up to 4 + it (plus the new support
// point at 4 + it).
__device__ int epa_wave(EpaShared& S, const Bodies& B, const Shape& sa, const pose& ta,
                        const WaveHull& ha, const Shape& sb, const pose& tb, const WaveHull& hb,
                        const f3 simplex[4], uint32_t iterations, float tol, f3& normal,
                        float& depth, uint32_t* face_out = nullptr, uint32_t* work = nullptr,
                        uint32_t* stop = nullptr)
{
    const uint32_t lane = lane_id();
    normal = {0.0f, 0.0f, 0.0f};
    depth = 0.0f;
    if (iterations == 0)
        return 0;
    // The caller may stage the simplex in S itself (simplex == nullptr).
    if (simplex && lane < 4) {
        S.vx[lane] = simplex[lane].x;
        S.vy[lane] = simplex[lane].y;
        S.vz[lane] = simplex[lane].z;
    }
    wave_sync();
    if (lane < 4) {
        // (0,1,2), (0,3,1), (0,2,3), (1,3,2) — epa.cpp:76-79.
        const uint32_t fa = lane == 3 ? 1u : 0u;
        const uint32_t fb = lane == 0 ? 1u : (lane == 1 ? 3u : (lane == 2 ? 2u : 3u));
        const uint32_t fc = lane == 0 ? 2u : (lane == 1 ? 1u : (lane == 2 ? 3u : 2u));
        make_face_lds(S, lane, fa, fb, fc);
    }
    wave_sync();
    uint32_t nf = 4, nv = 4;
    for (uint32_t it = 0; it < iterations; ++it) {
        if (nf == 0)
            { if (work) *work += it + 1; return 0; }
        // Long runs (mostly sphere-hull pairs, whose polytope approaches a
        // curved surface) bound the kernel's duration.  `work` counts the
        // iterations this wavefront has run, earlier pairs included: past
        // kLongEpa it takes issue priority over the wavefronts sharing its
        // SIMD, past kVeryLongEpa the highest (the caller resets it).
        if (work) {
            const uint32_t done = *work + it;
            if (done == kLongEpa || (it == 0 && done > kLongEpa && done < kVeryLongEpa))
                __builtin_amdgcn_s_setprio(2);
            if (done == kVeryLongEpa || (it == 0 && done > kVeryLongEpa))
                __builtin_amdgcn_s_setprio(3);
        } else if (it == kLongEpa) {
            __builtin_amdgcn_s_setprio(2);
        }
        // Closest face: first minimum of the face distances.
        float bd = 0.0f;
        uint32_t bi = kNoIndex;
        for (uint32_t f = lane; f < nf; f += 64) {
            const float d = S.fd[f];
            if (bi == kNoIndex || d < bd) {
                bd = d;
                bi = f;
            }
        }
        if (!wave_arg_best_lanes<false>(bd, bi, (nf + 63) / 64)) {
            wave_arg_best<false>(bd, bi);
            bi = uniform(bi);
        }
        const f3 cn{S.fnx[bi], S.fny[bi], S.fnz[bi]};
        const float cd = S.fd[bi];
        const f3 sup = sub(support_wave(B, sa, ta, cn, ha), support_wave(B, sb, tb, neg(cn), hb));
        if (dot(sup, cn) - cd <= tol) {
            normal = cn;
            depth = max_ref(cd, 0.0f);
            if (face_out)
                *face_out = uniform(S.fabc[bi]);
            { if (work) *work += it + 1; return 1; }
        }
        if (nv >= kMaxVertices)
            { if (work) *work += it + 1; return -1; }
        const uint32_t si = nv;
        if (lane == 0) {
            S.vx[si] = sup.x;
            S.vy[si] = sup.y;
            S.vz[si] = sup.z;
        }
        ++nv;
        // (The new vertex is read back only by make_face_lds, after the
        // synchronisations below.)
        // Visibility of every face, one ballot per chunk of 64; the visible
        // faces are listed in index order.  Chunks past nf are skipped.
        uint64_t vis[kFaceChunks] = {};
        uint32_t visible = 0;
#pragma unroll
        for (uint32_t ch = 0; ch < kFaceChunks; ++ch) {
            if (ch * 64 >= nf)
                break;
            const uint32_t f = ch * 64 + lane;
            bool v = false;
            if (f < nf) {
                const uint32_t fa = S.fabc[f] & 0xFFu;
                v = dot(f3{S.fnx[f], S.fny[f], S.fnz[f]}, sub(sup, lds_vertex(S, fa))) > tol;
            }
            vis[ch] = __ballot(v);
            if (v)
                S.vis_list[visible + (uint32_t)__popcll(vis[ch] & lanes_below(lane))] =
                    uint8_t(f);
            visible += (uint32_t)__popcll(vis[ch]);
        }
        wave_sync();
        const int horizon = horizon_wave(S, 3 * visible);
        if (horizon < 0) {
            if (work)
                *work += it + 1;
            if (stop)
                *stop = it | (nf << 16);
            return -1;
        }
        const uint32_t ne = uint32_t(horizon);
        if (ne == 0)
            { if (work) *work += it + 1; return 0; }
        const uint32_t keep = nf - visible;
        if (keep + ne > kMaxFaces) {
            if (work)
                *work += it + 1;
            if (stop)
                *stop = it | (nf << 16);
            return -1;
        }
        // Kept faces move down in order; every write lands at or below the
        // slot the writing lane read, and after everything already placed.
        uint32_t out = 0;
#pragma unroll
        for (uint32_t ch = 0; ch < kFaceChunks; ++ch) {
            if (ch * 64 >= nf)
                break;
            const uint32_t f = ch * 64 + lane;
            const bool valid = f < nf;
            uint32_t abc = 0;
            float nx = 0.0f, ny = 0.0f, nz = 0.0f, d = 0.0f;
            if (valid) {
                abc = S.fabc[f];
                nx = S.fnx[f];
                ny = S.fny[f];
                nz = S.fnz[f];
                d = S.fd[f];
            }
            const uint64_t kept = __ballot(valid) & ~vis[ch];
            wave_sync();
            if (valid && !((vis[ch] >> lane) & 1ull)) {
                const uint32_t pos = out + (uint32_t)__popcll(kept & lanes_below(lane));
                S.fabc[pos] = abc;
                S.fnx[pos] = nx;
                S.fny[pos] = ny;
                S.fnz[pos] = nz;
                S.fd[pos] = d;
            }
            out += (uint32_t)__popcll(kept);
            wave_sync();
        }
        for (uint32_t e = lane; e < ne; e += 64)
            make_face_lds(S, out + e, S.ea[e], S.eb[e], si);
        nf = out + ne;
        wave_sync();
    }
    if (work)
        *work += iterations;
    return 0;
}

// ------------------------------------------------------------ EPA past the bounds
// A pair whose polytope outgrows EpaShared (more than kMaxFaces faces or
// kMaxEdges horizon edges) is handed over, at the iteration where that
// happens, to a second kernel that continues computeEpa (epa.cpp:64-130)
// over distinct faces with multiplicities.  In FP32 the reference's list
// fills with copies of the same faces on near-degenerate pairs; all copies of
// a face share their visibility, so the run is fixed by:
//   - the closest face, the first minimum in list order;
//   - the horizon: for one unordered vertex pair the calls of addHorizonEdge
//     (epa.cpp:49-60) form a running sum s of +1 / -1 and the list ends with
//     |s| copies of the direction sign(s); s is a sum, independent of the
//     order of the calls, and a face with m copies contributes m per edge;
//   - the new faces, one per surviving directed edge, with |s| copies.
// Distinct faces are kept in the order of their first copies.  The list
// handed over is exact, so its faces (generation 0) are in known order; the
// faces of each later iteration follow all older ones, but their order among
// themselves is not tracked.  A closest-face minimum shared by faces of one
// later iteration with different normals is therefore undecided and reported
// (-2), as is a polytope of more than kContFaces distinct faces or a
// multiplicity past 2^32 - 1 (-1); every other run gives the reference's
// result.
constexpr uint32_t kContFaces = PHYSGPU_EPA_MAX_DISTINCT_FACES;

// Hand-over record (32-bit words): header, vertices (SoA), faces (SoA).
enum : uint32_t {
    kRecK = 0, kRecCls, kRecA, kRecB, kRecIt, kRecNf, kRecHeader = 8,
    kRecVx = kRecHeader, kRecVy = kRecVx + kMaxVertices, kRecVz = kRecVy + kMaxVertices,
    kRecFabc = kRecVz + kMaxVertices, kRecFnx = kRecFabc + kMaxFaces,
    kRecFny = kRecFnx + kMaxFaces, kRecFnz = kRecFny + kMaxFaces, kRecFd = kRecFnz + kMaxFaces,
    kRecEnd = kRecFd + kMaxFaces
};
static_assert(kRecEnd <= kEpaRecordWords, "hand-over record layout");

// One workgroup per handed-over run: kContThreads threads (one wavefront per
// SIMD of a CU), thread t holding faces kFacesPerThread t .. + 3, so block
// scans in thread order are scans in face order.
constexpr uint32_t kContThreads = 256;
constexpr uint32_t kContWaves = kContThreads / 64;
constexpr uint32_t kFacesPerThread = kContFaces / kContThreads;
static_assert(kFacesPerThread == 4, "four faces per thread (one float4 of distances)");
static_assert(kMaxFaces <= kContThreads, "a handed-over list is staged one face per thread");

struct ContShared {
    float vx[kMaxVertices], vy[kMaxVertices], vz[kMaxVertices];
    alignas(16) uint32_t fabc[kContFaces];  // a | b << 8 | c << 16
    alignas(16) float fnx[kContFaces];
    alignas(16) float fny[kContFaces];
    alignas(16) float fnz[kContFaces];
    alignas(16) float fd[kContFaces];
    alignas(16) uint32_t mult[kContFaces];
    alignas(16) uint8_t gen[kContFaces];  // 0: the handed-over faces; 1, 2, ...: later iterations
    // Signed running sum and first call index per unordered vertex pair:
    // zero / kNoIndex between iterations (each iteration's first calls reset
    // the pairs that iteration used).
    unsigned long long sum[kEdgeKeys];
    uint32_t first[kEdgeKeys];
    uint32_t vis_abc[kContFaces];  // visible faces in order
    uint32_t wsum[2][kContWaves];  // per-wavefront partials of the block scans (2 buffers)
    // Per wavefront: first minimum of the face distances its threads wrote
    // last, and flags (bit 0 a NaN distance, bit 1 past the storage bounds).
    float wbest_v[kContWaves];
    uint32_t wbest_i[kContWaves];
    uint32_t wflags[kContWaves];
};

// Edge e (0, 1, 2 = (a,b), (b,c), (c,a)) of a packed face.
__device__ __forceinline__ void face_edge(uint32_t abc, uint32_t e, uint32_t& x, uint32_t& y)
{
    const uint32_t v0 = abc & 0xFFu, v1 = (abc >> 8) & 0xFFu, v2 = (abc >> 16) & 0xFFu;
    x = e == 0 ? v0 : (e == 1 ? v1 : v2);
    y = e == 0 ? v1 : (e == 1 ? v2 : v0);
}

// Exclusive prefix sum, in thread order, of per-thread counts below 2^Bits
// over the workgroup, and its total; `any` (optional) receives the OR of a
// per-thread flag.  Every thread calls it; one barrier.  Consecutive calls
// alternate between the two partial buffers (`buf`), and any two calls on
// one buffer are separated by a barrier of the call between.
template <uint32_t Bits>
__device__ __forceinline__ uint32_t block_prefix(ContShared& S, uint32_t c, uint32_t& total,
                                                 uint32_t buf, bool flag = false,
                                                 bool* any = nullptr)
{
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t below = lanes_below(lane);
    uint32_t in_wave = 0, wave_total = 0;
#pragma unroll
    for (uint32_t b = 0; b < Bits; ++b) {
        const uint64_t m = __ballot((c >> b) & 1u);
        in_wave += (uint32_t)__popcll(m & below) << b;
        wave_total += (uint32_t)__popcll(m) << b;
    }
    const bool wave_flag = __ballot(flag) != 0;
    if (lane == 0)
        S.wsum[buf][w] = wave_total | (wave_flag ? 0x80000000u : 0u);
    __syncthreads();
    uint32_t off = 0, tot = 0, flags = 0;
#pragma unroll
    for (uint32_t q = 0; q < kContWaves; ++q) {
        const uint32_t x = S.wsum[buf][q];
        flags |= x;
        off += q < w ? x & 0x7FFFFFFFu : 0u;
        tot += x & 0x7FFFFFFFu;
    }
    total = tot;
    if (any)
        *any = (flags & 0x80000000u) != 0;
    return off + in_wave;
}

// Publishes this wavefront's first minimum (v, i) of the face distances its
// threads wrote (i == kNoIndex: none) and its NaN / storage-bound flags; the
// caller's next barrier makes them visible to closest_face.
__device__ __forceinline__ void publish_best(ContShared& S, float v, uint32_t i, bool nan,
                                             bool bound)
{
    const uint32_t w = threadIdx.x >> 6;
    const bool wave_nan = __ballot(nan) != 0, wave_bound = __ballot(bound) != 0;
    wave_min_first(v, i);
    if (lane_id() == 0) {
        S.wbest_v[w] = v;
        S.wbest_i[w] = i;
        S.wflags[w] = (wave_nan ? 1u : 0u) | (wave_bound ? 2u : 0u);
    }
}

__device__ __forceinline__ bool published_bound(const ContShared& S)
{
    uint32_t flags = 0;
#pragma unroll
    for (uint32_t q = 0; q < kContWaves; ++q)
        flags |= S.wflags[q];
    return (flags & 2u) != 0;
}

// Closest face from the published minima: the first minimum of S.fd[0, nf)
// in index order, as the reference's sequential scan (std::min_element)
// picks it; with a NaN distance among the faces, that scan itself.  Every
// thread gets the index.
__device__ __forceinline__ uint32_t closest_face(const ContShared& S, uint32_t nf)
{
    float bv = 0.0f;
    uint32_t bi = kNoIndex, flags = 0;
#pragma unroll
    for (uint32_t q = 0; q < kContWaves; ++q) {
        flags |= S.wflags[q];
        arg_take<false>(bv, bi, S.wbest_v[q], S.wbest_i[q]);
    }
    if (flags & 1u) {
        uint32_t best = 0;
        for (uint32_t f = 1; f < nf; ++f)
            if (S.fd[f] < S.fd[best])
                best = f;
        bi = best;
    }
    return bi;
}

// Iterations [first, iterations) over the nf distinct faces in S (first-copy
// order, multiplicities in S.mult, generation 0) and the vertices up to
// 4 + first, with the pair sums cleared and the closest-face minima of those
// faces published (publish_best, then a barrier).  Returns 1 converged, 0 the
// reference's failure, -1 storage bound, -2 undecided closest face; normal /
// depth uniform across the block.  Every thread of the workgroup calls it
// (ha, hb hold both shapes' points per lane).  Barriers per iteration: the
// scan of the visible faces (which also ORs the undecided test), the end of
// the horizon, one scan per kContThreads horizon calls, and the end of the
// writes (which also publishes the next closest face).
__device__ int epa_multi_block(ContShared& S, const Bodies& B, const Shape& sa, const pose& ta,
                               const WaveHull& ha, const Shape& sb, const pose& tb,
                               const WaveHull& hb, uint32_t first, uint32_t nf,
                               uint32_t iterations, float tol, f3& normal, float& depth)
{
    const uint32_t t = threadIdx.x, f0 = kFacesPerThread * t;
    normal = {0.0f, 0.0f, 0.0f};
    depth = 0.0f;
    uint32_t nv = 4 + first;
    uint32_t gen = 1;
    uint32_t buf = 0;
    for (uint32_t it = first; it < iterations; ++it) {
        if (nf == 0)
            return 0;
        const uint32_t bi = closest_face(S, nf);
        const f3 cn{S.fnx[bi], S.fny[bi], S.fnz[bi]};
        const float cd = S.fd[bi];
        const uint32_t g0 = S.gen[bi];
        const uint32_t si = nv;
        // This thread's faces (one 16-byte read per array).
        const uint4 abc4 = reinterpret_cast<const uint4*>(S.fabc)[t];
        const float4 nx4 = reinterpret_cast<const float4*>(S.fnx)[t];
        const float4 ny4 = reinterpret_cast<const float4*>(S.fny)[t];
        const float4 nz4 = reinterpret_cast<const float4*>(S.fnz)[t];
        const float4 d4 = reinterpret_cast<const float4*>(S.fd)[t];
        const uint4 m4 = reinterpret_cast<const uint4*>(S.mult)[t];
        const uint32_t g4 = reinterpret_cast<const uint32_t*>(S.gen)[t];
        const uint32_t abc[kFacesPerThread] = {abc4.x, abc4.y, abc4.z, abc4.w};
        const uint32_t m[kFacesPerThread] = {m4.x, m4.y, m4.z, m4.w};
        const uint8_t g[kFacesPerThread] = {uint8_t(g4), uint8_t(g4 >> 8), uint8_t(g4 >> 16),
                                            uint8_t(g4 >> 24)};
        const float nx[kFacesPerThread] = {nx4.x, nx4.y, nx4.z, nx4.w};
        const float ny[kFacesPerThread] = {ny4.x, ny4.y, ny4.z, ny4.w};
        const float nz[kFacesPerThread] = {nz4.x, nz4.y, nz4.z, nz4.w};
        const float d[kFacesPerThread] = {d4.x, d4.y, d4.z, d4.w};
        // Every wavefront evaluates the support maps (same bits), so none
        // waits for another's.
        const f3 sup = sub(support_wave(B, sa, ta, cn, ha), support_wave(B, sb, tb, neg(cn), hb));
        if (t == 0) {
            S.vx[si] = sup.x;
            S.vy[si] = sup.y;
            S.vz[si] = sup.z;
        }
        ++nv;
        // Whether one of this thread's faces ties the closest face within its
        // iteration (undecided), and which are visible.
        bool tie = false;
        bool vis[kFacesPerThread];
        uint32_t nvis = 0;
#pragma unroll
        for (uint32_t j = 0; j < kFacesPerThread; ++j) {
            const uint32_t f = f0 + j;
            tie = tie || (g0 != 0 && f < nf && f > bi && g[j] == g0 && d[j] == cd &&
                          (nx[j] != cn.x || ny[j] != cn.y || nz[j] != cn.z));
            vis[j] = f < nf &&
                     dot(f3{nx[j], ny[j], nz[j]}, sub(sup, lds_vertex(S, abc[j] & 0xFFu))) > tol;
            nvis += vis[j] ? 1u : 0u;
        }
        const bool converged = dot(sup, cn) - cd <= tol;
        // The pair sums do not depend on the order of the calls: they go in
        // before the scan that orders the visible faces.
        if (!converged) {
#pragma unroll
            for (uint32_t j = 0; j < kFacesPerThread; ++j)
                if (vis[j])
#pragma unroll
                    for (uint32_t e = 0; e < 3; ++e) {
                        uint32_t x, y;
                        face_edge(abc[j], e, x, y);
                        const long long delta = x < y ? (long long)m[j] : -(long long)m[j];
                        atomicAdd(&S.sum[edge_key(x, y)], (unsigned long long)delta);
                    }
        }
        uint32_t visible = 0;
        bool undecided = false;
        const uint32_t before = block_prefix<3>(S, nvis, visible, buf, tie, &undecided);
        buf ^= 1u;
        if (undecided)
            return -2;
        if (converged) {
            normal = cn;
            depth = max_ref(cd, 0.0f);
            return 1;
        }
        // First call of every pair (call 3 r + e is edge e of visible face
        // r; this thread's calls are 3 before ... 3 (before + nvis) - 1, in
        // order); kept faces move down in order, their first minimum taken
        // on the way.
        float bv = 0.0f;
        uint32_t bidx = kNoIndex;
        bool nan = false;
        uint32_t r = 0;
#pragma unroll
        for (uint32_t j = 0; j < kFacesPerThread; ++j) {
            const uint32_t f = f0 + j;
            if (vis[j]) {
                S.vis_abc[before + r] = abc[j];
#pragma unroll
                for (uint32_t e = 0; e < 3; ++e) {
                    uint32_t x, y;
                    face_edge(abc[j], e, x, y);
                    atomicMin(&S.first[edge_key(x, y)], 3 * (before + r) + e);
                }
                ++r;
            } else if (f < nf) {
                const uint32_t pos = f - before - r;
                S.fabc[pos] = abc[j];
                S.fnx[pos] = nx[j];
                S.fny[pos] = ny[j];
                S.fnz[pos] = nz[j];
                S.fd[pos] = d[j];
                S.mult[pos] = m[j];
                S.gen[pos] = g[j];
                nan = nan || d[j] != d[j];
                if (bidx == kNoIndex || d[j] < bv) {
                    bv = d[j];
                    bidx = pos;
                }
            }
        }
        __syncthreads();
        // New faces: one per unordered vertex pair with a nonzero sum, made
        // at the pair's first call, in call order; calls taken kContThreads at
        // a time.  The thread of a pair's first call also clears the pair
        // (the pair's other calls only compare its first call with theirs).
        const uint32_t keep = nf - visible, calls = 3 * visible;
        uint32_t ne = 0;
        bool bound = false;
        for (uint32_t base = 0; base < calls; base += kContThreads) {
            const uint32_t call = base + t;
            bool emit = false;
            uint32_t x = 0, y = 0;
            long long sk = 0;
            if (call < calls) {
                face_edge(S.vis_abc[call / 3], call % 3, x, y);
                const uint32_t key = edge_key(x, y);
                if (S.first[key] == call) {
                    sk = (long long)S.sum[key];
                    emit = sk != 0;
                    S.sum[key] = 0ull;
                    S.first[key] = kNoIndex;
                }
            }
            uint32_t round = 0;
            const uint32_t pos = keep + ne + block_prefix<1>(S, emit ? 1u : 0u, round, buf);
            buf ^= 1u;
            if (emit) {
                const uint32_t lo = x < y ? x : y, hi = x < y ? y : x;
                const unsigned long long mm = (unsigned long long)(sk > 0 ? sk : -sk);
                if (pos < kContFaces && mm <= 0xFFFFFFFFull) {
                    const float dn = make_face_lds(S, pos, sk > 0 ? lo : hi, sk > 0 ? hi : lo, si);
                    S.mult[pos] = uint32_t(mm);
                    S.gen[pos] = uint8_t(gen);
                    nan = nan || dn != dn;
                    if (bidx == kNoIndex || dn < bv) {
                        bv = dn;
                        bidx = pos;
                    }
                } else {
                    bound = true;
                }
            }
            ne += round;
        }
        if (ne == 0)
            return 0;
        publish_best(S, bv, bidx, nan, bound);
        __syncthreads();
        if (published_bound(S))
            return -1;
        nf = keep + ne;
        ++gen;
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

constexpr uint32_t kClassifyThreads = 1024;
constexpr uint32_t kClassifyGrid = 256;

// Shape-pair classification and bucketing.  Counts are aggregated per
// workgroup of 1024 candidates in LDS, so each class costs one global atomic
// per workgroup and round.
__global__ __launch_bounds__(kClassifyThreads) void k_classify(Bodies B, World W,
                                                               const uint64_t* __restrict__ pairs,
                                                               Counters* c,
                                                               uint32_t* __restrict__ buckets,
                                                               uint64_t cap,
                                                               uint32_t* __restrict__ hit,
                                                               uint64_t* __restrict__ epa_partials,
                                                               uint32_t* __restrict__ epa_epoch)
{
    // A new epoch for the EPA hand-over flags of this narrow phase.
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *epa_epoch += 1u;
    // The EPA kernels that follow add their counts into these partial slots.
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kEpaPartialSlots * kEpaPartials;
         i += gridDim.x * blockDim.x)
        epa_partials[(i / kEpaPartials) * kMaxPartialBlocks + i % kEpaPartials] = 0;
    constexpr uint32_t kWavesPerBlock = kClassifyThreads / 64;
    __shared__ uint32_t wave_count[kWavesPerBlock][kClasses];
    __shared__ uint32_t block_base[kClasses];
    // Pairs held: the broad phase's count clamped to the capacity (written
    // here for the kernels after the narrow phase; the canonical sort on the
    // side stream clamps it itself).
    const unsigned long long emitted = c->pair_count;
    const uint32_t count = uint32_t(emitted < cap ? emitted : cap);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        c->sort_count = count;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (count + stride - 1) / stride;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
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
                cls = pair_class(load_type(B, a), load_type(B, b));
        }
        uint64_t mine = 0;
#pragma unroll
        for (uint32_t q = 0; q < kClasses; ++q) {
            const uint64_t m = __ballot(cls == q);
            if (cls == q)
                mine = m;
            if (lane == 0)
                wave_count[wave][q] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (threadIdx.x < kClasses) {
            const uint32_t q = threadIdx.x;
            uint32_t run = 0;
            for (uint32_t w = 0; w < kWavesPerBlock; ++w) {
                const uint32_t v = wave_count[w][q];
                wave_count[w][q] = run;
                run += v;
            }
            block_base[q] = run ? atomicAdd(&c->class_count[q], run) : 0u;
        }
        __syncthreads();
        if (cls < kClasses) {
            const uint32_t slot = block_base[cls] + wave_count[wave][cls] +
                                  (uint32_t)__popcll(mine & lanes_below(lane));
            buckets[cls * cap + slot] = k;
        }
        __syncthreads();
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
                                                  uint32_t* __restrict__ hit,
                                                  uint64_t* __restrict__ contacts_part)
{
    const uint32_t total = c->class_count[kClass];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (total + stride - 1) / stride;
    uint32_t hits = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t t = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool h = false;
        const uint32_t cls = kClass;
        if (t < total) {
            const uint32_t k = buckets[cls * cap + t];
            const uint64_t pr = pairs[k];
            const uint32_t a = uint32_t(pr >> 32), b = uint32_t(pr);
            const pose ta = load_pose(B, a), tb = load_pose(B, b);
            Contact ct;
            if (kClass == PHYSGPU_PAIR_SPHERE_SPHERE) {
                h = sphere_sphere(load_data(B, a).x, ta, load_data(B, b).x, tb, ct);
            } else if (kClass == PHYSGPU_PAIR_SPHERE_BOX) {
                if (load_type(B, a) == PHYSGPU_SPHERE) {
                    h = sphere_box(load_data(B, a).x, ta, load_data(B, b), tb, ct);
                } else {
                    // Box-sphere: swap, then negate the normal and swap anchors (:353-365).
                    h = sphere_box(load_data(B, b).x, tb, load_data(B, a), ta, ct);
                    if (h) {
                        ct.normal = neg(ct.normal);
                        const f3 tmp = ct.a[0];
                        ct.a[0] = ct.b[0];
                        ct.b[0] = tmp;
                    }
                }
            } else if (kClass == PHYSGPU_PAIR_BOX_BOX) {
                h = box_box(load_data(B, a), ta, load_data(B, b), tb, ct);
            }
            if (h) {
                store_contact(ct, k, a, b, cls, W, B.n, slots);
                hit[k] = 1;
                ++hits;
            }
        }
    }
    block_partial(hits, contacts_part);
}

__device__ __forceinline__ Shape load_shape(const Bodies& B, uint32_t i)
{
    Shape s{load_type(B, i), load_data(B, i), 0, 0};
    if (s.type == PHYSGPU_HULL)
        hull_range(B.shape, i, s.first, s.count);
    return s;
}

// GJK (intersectGjk, gjk.cpp:128-158) on 16 lanes (one 16-lane DPP row).
// Each lane holds one point of each hull, so a hull support map is one dot
// product and a 4-step row arg-max (first maximum in index order, as the
// sequential strict-> scan) instead of a loop of dependent loads; the simplex
// logic runs replicated in the row's lanes.
constexpr uint32_t kGjkLanes = 16;

// A hull's points (at most 16) or a box's 8 corners held one per row lane,
// in collider space (p) and in world space (w = to_world(p): the function the
// support map applies to the chosen point, so the same bits).  Box corner c
// has the +extent on the axes of its set bits (bit 0 x, bit 1 y, bit 2 z).
struct RowHull {
    f3 p, w;
    bool cached;  // hull of at most 16 points
    bool box;
};

__device__ __forceinline__ RowHull row_hull(const Bodies& B, const Shape& s, const pose& t,
                                            uint32_t rl)
{
    RowHull h{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f},
              s.type == PHYSGPU_HULL && s.count <= kGjkLanes, s.type == PHYSGPU_BOX};
    if (h.cached && rl < s.count) {
        const uint32_t p = s.first + rl, nh = B.nh;
        h.p = {B.hpts[p], B.hpts[nh + p], B.hpts[2 * nh + p]};
        h.w = to_world(h.p, t);
    } else if (h.box && rl < 8) {
        h.p = {(rl & 1u) ? s.data.x : -s.data.x, (rl & 2u) ? s.data.y : -s.data.y,
               (rl & 4u) ? s.data.z : -s.data.z};
        h.w = to_world(h.p, t);
    }
    return h;
}

__device__ __forceinline__ f3 row_point(const f3& v, uint32_t row_base, uint32_t src)
{
    const int at = int(row_base + src);
    return {__shfl(v.x, at), __shfl(v.y, at), __shfl(v.z, at)};
}

// Support map evaluated by a row (same result in its 16 lanes).  All 16
// lanes of a row call it together.
__device__ f3 support_row(const Bodies& B, const Shape& s, const pose& t, f3 d, const RowHull& h,
                          uint32_t rl, uint32_t row_base)
{
    if (s.type != PHYSGPU_HULL) {
        if (h.box) {
            // boxSupportPoint's corner (narrowphase.cpp:27-35).
            const f3 l = rotate(conj(t.q), d);
            const uint32_t c = (l.x >= 0.0f ? 1u : 0u) | (l.y >= 0.0f ? 2u : 0u) |
                               (l.z >= 0.0f ? 4u : 0u);
            return row_point(h.w, row_base, c);
        }
        return support(B, s, t, d);
    }
    const f3 l = rotate(conj(t.q), d);
    const uint32_t nh = B.nh;
    float bv = 0.0f;
    uint32_t bi = kNoIndex;
    if (h.cached) {
        if (rl < s.count) {
            bv = dot(h.p, l);
            bi = rl;
        }
    } else {
        for (uint32_t k = rl; k < s.count; k += kGjkLanes) {
            const uint32_t p = s.first + k;
            const float v = dot(f3{B.hpts[p], B.hpts[nh + p], B.hpts[2 * nh + p]}, l);
            if (bi == kNoIndex || v > bv) {
                bv = v;
                bi = k;
            }
        }
    }
    if (!row_arg_best_lanes<true>(bv, bi, (s.count + kGjkLanes - 1) / kGjkLanes, row_base))
        row_arg_best<true>(bv, bi);
    if (bi == kNoIndex)
        bi = 0;  // no points: the first slot, as the sequential scan starts there
    if (h.cached)
        return row_point(h.w, row_base, bi);
    const uint32_t p = s.first + bi;
    return to_world(f3{B.hpts[p], B.hpts[nh + p], B.hpts[2 * nh + p]}, t);
}

// Hull pairs: GJK (the row code above) and, for an intersecting pair, EPA
// by the whole wavefront; the contact point is the midpoint of the two
// deepest support points along the EPA normal (A toward B).  A work item is
// one sphere-hull candidate (the four rows on the same pair, so its possibly
// long EPA run starts at once) or four candidates, one per row, whose EPA
// runs follow one another.  Each workgroup is one wavefront and takes item
// blockIdx.x (then every gridDim.x-th after it): with at least as many
// workgroups as items the dispatcher hands them out in order as wavefront
// slots free up, a work queue with no atomics.  The sphere-hull bucket comes
// first, so its long runs start first, and a pair's EPA follows its own GJK
// with no grid-wide wait in between.  Counts go to the EPA partial slots (zeroed by
// k_classify) with one atomic add per nonzero count.
constexpr uint32_t kEpaBlock = 64;

__device__ __forceinline__ void epa_partial_add(uint64_t* slot, uint32_t v)
{
    if (v)
        atomicAdd(reinterpret_cast<unsigned long long*>(slot + (blockIdx.x & (kEpaPartials - 1))),
                  (unsigned long long)v);
}

constexpr uint32_t kRowsPerWave = 64 / kGjkLanes;

// Hull candidate t of the class 3, 4, 5 buckets taken in that order.
struct Candidate {
    uint32_t cls, k, a, b;
};

__device__ __forceinline__ Candidate candidate(uint32_t t, uint32_t c3, uint32_t c4,
                                               const uint32_t* __restrict__ buckets, uint64_t cap,
                                               const uint64_t* __restrict__ pairs)
{
    uint32_t cls = 3, q = t;
    if (q >= c3 + c4) {
        cls = 5;
        q -= c3 + c4;
    } else if (q >= c3) {
        cls = 4;
        q -= c3;
    }
    const uint32_t k = buckets[cls * cap + q];
    const uint64_t pr = pairs[k];
    return {cls, k, uint32_t(pr >> 32), uint32_t(pr)};
}

// Hand-over record of a run stopped by the LDS bounds (epa_wave's *stop).
// Kept out of line: it runs on a few pairs per million and its registers
// would otherwise count against the EPA loop's.
__device__ __attribute__((noinline)) void write_handover(const EpaShared& S, uint32_t* R,
                                                         uint32_t k, uint32_t cls, uint32_t a,
                                                         uint32_t b, uint32_t stop)
{
    const uint32_t lane = lane_id();
    const uint32_t it = stop & 0xFFFFu, nf = stop >> 16, nv = 4 + it;
    if (lane == 0) {
        R[kRecK] = k;
        R[kRecCls] = cls;
        R[kRecA] = a;
        R[kRecB] = b;
        R[kRecIt] = it;
        R[kRecNf] = nf;
    }
    for (uint32_t i = lane; i < nv; i += 64) {
        R[kRecVx + i] = __float_as_uint(S.vx[i]);
        R[kRecVy + i] = __float_as_uint(S.vy[i]);
        R[kRecVz + i] = __float_as_uint(S.vz[i]);
    }
    for (uint32_t f = lane; f < nf; f += 64) {
        R[kRecFabc + f] = S.fabc[f];
        R[kRecFnx + f] = __float_as_uint(S.fnx[f]);
        R[kRecFny + f] = __float_as_uint(S.fny[f]);
        R[kRecFnz + f] = __float_as_uint(S.fnz[f]);
        R[kRecFd + f] = __float_as_uint(S.fd[f]);
    }
}

__global__ __launch_bounds__(kEpaBlock) __attribute__((amdgpu_waves_per_eu(4)))
void k_gjk_epa(Bodies B, World W, const uint64_t* __restrict__ pairs, Counters* c,
               const uint32_t* __restrict__ buckets, uint64_t cap,
               physgpu_narrowphase_options opt, physgpu_manifold* __restrict__ slots,
               uint32_t* __restrict__ hit, uint64_t* __restrict__ partials, bool solo_sphere_hull,
               uint32_t* __restrict__ records, uint32_t record_cap,
               uint64_t* __restrict__ unresolved, uint32_t* __restrict__ ready,
               const uint32_t* __restrict__ epa_epoch)
{
    const uint32_t epoch = *epa_epoch;
    // Tells k_epa_early that this kernel runs beside it.
    if (blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(&ready[2 * record_cap + 1], epoch, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    __shared__ EpaShared S;
    // GJK simplices (4 points x 3) and candidate positions of the rows that
    // found an intersection, for the EPA runs that follow.
    __shared__ float stage_p[kRowsPerWave][12];
    __shared__ uint32_t stage_t[kRowsPerWave];
    const uint32_t lane = lane_id(), rl = lane % kGjkLanes, row_base = lane - rl;
    const uint32_t row = lane / kGjkLanes;
    const uint32_t c3 = c->class_count[3], c4 = c->class_count[4];
    const uint32_t total = c3 + c4 + c->class_count[5];
    const uint32_t solo_items = solo_sphere_hull ? c3 : 0;
    const uint32_t items = solo_items + (total - solo_items + kRowsPerWave - 1) / kRowsPerWave;
    uint32_t contacts3 = 0, contacts4 = 0, contacts5 = 0;
    uint32_t failures = 0, overflows = 0;
    uint32_t work = 0;  // EPA iterations of this wavefront so far
    uint32_t items_done = 0;
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x, ++items_done) {
        // Sphere-hull items are one candidate (all four rows on it) when
        // solo_sphere_hull is set; the other candidates go four to an item,
        // one per row, and run their EPAs in turn.
        const bool solo = item < solo_items;
        const uint32_t tr = solo ? item : solo_items + kRowsPerWave * (item - solo_items) + row;
        const bool valid = tr < total;
        Simplex sx;
        bool intersect = false;
        {
            const uint32_t t = valid ? tr : total - 1;
            const Candidate cd = candidate(t, c3, c4, buckets, cap, pairs);
            const Shape sa = load_shape(B, cd.a), sb = load_shape(B, cd.b);
            const pose ta = load_pose(B, cd.a), tb = load_pose(B, cd.b);
            const RowHull ra = row_hull(B, sa, ta, rl), rb = row_hull(B, sb, tb, rl);
            sx.size = 0;
            // A row past the last candidate stops at once (zero direction).
            f3 dir{valid ? 1.0f : 0.0f, 0.0f, 0.0f};
            for (uint32_t it = 0; it < opt.gjk_max_iterations; ++it) {
                if (dot(dir, dir) < 1e-12f)
                    break;
                const f3 pa = support_row(B, sa, ta, dir, ra, rl, row_base);
                const f3 pb = support_row(B, sb, tb, neg(dir), rb, rl, row_base);
                const f3 p = sub(pa, pb);
                if (dot(p, dir) <= 0.0f)
                    break;
                // Shift the simplex up by one (constant indices keep it in
                // registers).
                if (sx.size >= 3)
                    sx.p[3] = sx.p[2];
                if (sx.size >= 2)
                    sx.p[2] = sx.p[1];
                if (sx.size >= 1)
                    sx.p[1] = sx.p[0];
                sx.p[0] = p;
                sx.size = min(sx.size + 1, 4u);
                if (gjk_contains(sx, dir)) {
                    intersect = true;
                    break;
                }
            }
        }
        // Bit 16 r: row r intersects.
        uint64_t rows_hit = __ballot(intersect && valid && rl == 0);
        if (solo)
            rows_hit &= 1ull;
        if (!rows_hit)
            continue;
        if (intersect && valid && rl == 0 && (!solo || row == 0)) {
#pragma unroll
            for (uint32_t v = 0; v < 4; ++v) {
                stage_p[row][3 * v] = sx.p[v].x;
                stage_p[row][3 * v + 1] = sx.p[v].y;
                stage_p[row][3 * v + 2] = sx.p[v].z;
            }
            stage_t[row] = tr;
        }
        wave_sync();
        while (rows_hit) {
            const uint32_t r = uint32_t(__ffsll((unsigned long long)rows_hit) - 1) / kGjkLanes;
            rows_hit &= rows_hit - 1;
            const Candidate cd = candidate(uniform(stage_t[r]), c3, c4, buckets, cap, pairs);
            const uint32_t a = cd.a, b = cd.b, k = cd.k, cls = cd.cls;
            const Shape sa = load_shape(B, a), sb = load_shape(B, b);
            const pose ta = load_pose(B, a), tb = load_pose(B, b);
            if (lane < 4) {
                S.vx[lane] = stage_p[r][3 * lane];
                S.vy[lane] = stage_p[r][3 * lane + 1];
                S.vz[lane] = stage_p[r][3 * lane + 2];
            }
            f3 n;
            float depth;
            const WaveHull ha = wave_hull(B, sa, ta), hb = wave_hull(B, sb, tb);
            uint32_t stop = 0;
            const int st = epa_wave(S, B, sa, ta, ha, sb, tb, hb, nullptr,
                                    opt.epa_max_iterations, opt.epa_tolerance, n, depth, nullptr,
                                    &work, &stop);
            if (st == 1) {
                const f3 w = scale(add(support_wave(B, sa, ta, n, ha),
                                       support_wave(B, sb, tb, neg(n), hb)),
                                   0.5f);
                if (lane == 0) {
                    Contact ct;
                    ct.normal = n;
                    ct.count = 1;
                    ct.depth[0] = depth;
                    set_anchors(ct, 0, w, ta, tb);
                    store_contact(ct, k, a, b, cls, W, B.n, slots);
                    hit[k] = 1;
                }
                contacts3 += cls == 3 ? 1u : 0u;
                contacts4 += cls == 4 ? 1u : 0u;
                contacts5 += cls == 5 ? 1u : 0u;
            } else if (st == 0) {
                ++failures;
            } else {
                // Past the LDS bounds: hand the exact state of this iteration
                // over to k_epa_continue (or, with no record left, report the
                // pair as unresolved).
                ++overflows;
                uint32_t slot = 0;
                if (lane == 0)
                    slot = atomicAdd(&c->epa_continued, 1u);
                slot = uniform(slot);
                if (slot < record_cap) {
                    write_handover(S, records + size_t(slot) * kEpaRecordWords, k, cls, a, b,
                                   stop);
                    // Publish the record to the consumers (every lane's
                    // writes first, then the flag with release semantics).
                    __threadfence();
                    wave_sync();
                    if (lane == 0)
                        __hip_atomic_store(&ready[slot], epoch, __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_AGENT);
                } else if (lane == 0) {
                    const uint32_t u = atomicAdd(&c->epa_unresolved, 1u);
                    if (u < record_cap)
                        unresolved[u] = (uint64_t(a) << 32) | b;
                }
            }
            wave_sync();
        }
    }
    __builtin_amdgcn_s_setprio(0);
    // Counts are wave-uniform: lane 0 adds the wavefront's share.
    if (lane == 0) {
        epa_partial_add(partials + size_t(kSlotContacts + 3) * kMaxPartialBlocks, contacts3);
        epa_partial_add(partials + size_t(kSlotContacts + 4) * kMaxPartialBlocks, contacts4);
        epa_partial_add(partials + size_t(kSlotContacts + 5) * kMaxPartialBlocks, contacts5);
        epa_partial_add(partials + size_t(kSlotGjkFailures) * kMaxPartialBlocks, failures);
        epa_partial_add(partials + size_t(kSlotEpaOverflows) * kMaxPartialBlocks, overflows);
        // Items finished (relaxed: k_epa_early only uses the count to stop
        // looking; a record it misses is k_epa_continue's).
        if (items_done)
            __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(
                                       partials + size_t(kSlotEpaItemsDone) * kMaxPartialBlocks +
                                       (blockIdx.x % kEpaDoneSlots) * kEpaDoneStride),
                                   (unsigned long long)items_done, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One handed-over run (record R), continued over distinct faces with
// multiplicities (epa_multi_block) by the whole workgroup: the exact list
// handed over becomes distinct faces in first-copy order with their copy
// counts.  Results are stored like k_gjk_epa's; undecided runs are listed as
// unresolved.
__device__ void continue_record(ContShared& S, const Bodies& B, const World& W, Counters* c,
                                const uint32_t* __restrict__ R, uint32_t record_cap,
                                const physgpu_narrowphase_options& opt,
                                physgpu_manifold* __restrict__ slots, uint32_t* __restrict__ hit,
                                uint64_t* __restrict__ partials,
                                uint64_t* __restrict__ unresolved)
{
    const uint32_t t = threadIdx.x;
    const uint32_t k = R[kRecK], cls = R[kRecCls], a = R[kRecA], b = R[kRecB];
    const uint32_t it0 = R[kRecIt], nf0 = R[kRecNf];
    const Shape sa = load_shape(B, a), sb = load_shape(B, b);
    const pose ta = load_pose(B, a), tb = load_pose(B, b);
    const WaveHull ha = wave_hull(B, sa, ta), hb = wave_hull(B, sb, tb);
    if (t < 4 + it0) {
        S.vx[t] = __uint_as_float(R[kRecVx + t]);
        S.vy[t] = __uint_as_float(R[kRecVy + t]);
        S.vz[t] = __uint_as_float(R[kRecVz + t]);
    }
    if (t < nf0) {  // nf0 <= kMaxFaces = kContThreads
        S.fabc[t] = R[kRecFabc + t];
        S.mult[t] = 0;
    }
    __syncthreads();
    // First copy of every face (a face's copies have equal a, b, c and hence
    // equal normal and distance bits); copies counted on it.
    bool lead = false;
    if (t < nf0) {
        const uint32_t abc = S.fabc[t];
        uint32_t firstc = t;
        for (uint32_t g = 0; g < t; ++g)
            if (S.fabc[g] == abc) {
                firstc = g;
                break;
            }
        lead = firstc == t;
        atomicAdd(&S.mult[firstc], 1u);
    }
    __syncthreads();
    uint32_t abc = 0, m = 0;
    float nx = 0.0f, ny = 0.0f, nz = 0.0f, d = 0.0f;
    if (lead) {
        abc = R[kRecFabc + t];
        nx = __uint_as_float(R[kRecFnx + t]);
        ny = __uint_as_float(R[kRecFny + t]);
        nz = __uint_as_float(R[kRecFnz + t]);
        d = __uint_as_float(R[kRecFd + t]);
        m = S.mult[t];
    }
    uint32_t nf = 0;
    const uint32_t pos = block_prefix<1>(S, lead ? 1u : 0u, nf, 0);
    if (lead) {
        S.fabc[pos] = abc;
        S.fnx[pos] = nx;
        S.fny[pos] = ny;
        S.fnz[pos] = nz;
        S.fd[pos] = d;
        S.mult[pos] = m;
        S.gen[pos] = 0;
    }
    for (uint32_t q = t; q < kEdgeKeys; q += kContThreads) {
        S.sum[q] = 0ull;
        S.first[q] = kNoIndex;
    }
    // Each lead is its own first minimum (one face per thread).
    publish_best(S, d, lead ? pos : kNoIndex, lead && d != d, false);
    __syncthreads();
    f3 n;
    float depth;
    const int st = epa_multi_block(S, B, sa, ta, ha, sb, tb, hb, it0, nf, opt.epa_max_iterations,
                                   opt.epa_tolerance, n, depth);
    if (t < 64) {  // wavefront 0 stores the result
        if (st == 1) {
            const f3 w = scale(add(support_wave(B, sa, ta, n, ha),
                                   support_wave(B, sb, tb, neg(n), hb)),
                               0.5f);
            if (t == 0) {
                Contact ct;
                ct.normal = n;
                ct.count = 1;
                ct.depth[0] = depth;
                set_anchors(ct, 0, w, ta, tb);
                store_contact(ct, k, a, b, cls, W, B.n, slots);
                hit[k] = 1;
                atomicAdd(reinterpret_cast<unsigned long long*>(
                              partials + size_t(kSlotContacts + cls) * kMaxPartialBlocks),
                          1ull);
            }
        } else if (st == 0) {
            if (t == 0)
                atomicAdd(reinterpret_cast<unsigned long long*>(
                              partials + size_t(kSlotGjkFailures) * kMaxPartialBlocks),
                          1ull);
        } else if (t == 0) {
            const uint32_t u = atomicAdd(&c->epa_unresolved, 1u);
            if (u < record_cap)
                unresolved[u] = (uint64_t(a) << 32) | b;
        }
    }
    __syncthreads();
}

// After k_gjk_epa: every handed-over run that k_epa_early has not claimed
// (workgroup g takes records g, g + gridDim.x, ...).  claim[record_cap] is
// the word that tells k_epa_early this kernel has started.
__global__ __launch_bounds__(kContThreads) void k_epa_continue(
    Bodies B, World W, Counters* c, const uint32_t* __restrict__ records, uint32_t record_cap,
    physgpu_narrowphase_options opt, physgpu_manifold* __restrict__ slots,
    uint32_t* __restrict__ hit, uint64_t* __restrict__ partials, uint64_t* __restrict__ unresolved,
    uint32_t* __restrict__ claim, const uint32_t* __restrict__ epa_epoch)
{
    __shared__ ContShared S;
    __shared__ uint32_t won;
    const uint32_t epoch = *epa_epoch;
    // k_gjk_epa has finished: k_epa_early stops taking runs.
    if (blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(&claim[record_cap], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t count = min(c->epa_continued, record_cap);
    for (uint32_t r = blockIdx.x; r < count; r += gridDim.x) {
        if (threadIdx.x == 0)
            won = atomicExch(&claim[r], epoch) != epoch ? 1u : 0u;
        __syncthreads();
        if (won)
            continue_record(S, B, W, c, records + size_t(r) * kEpaRecordWords, record_cap, opt,
                            slots, hit, partials, unresolved);
        __syncthreads();
    }
}

// While k_gjk_epa still runs (on another stream): take handed-over runs as
// they are published (ready flag of this epoch, claim word).  Stops when
// every GJK + EPA work item has finished and no published run is left to
// take, when k_epa_continue has started (it takes every run not taken here),
// when the wait passes kEpaEarlyPollTicks, or when k_gjk_epa has not started
// within kEpaEarlyStartTicks (the two streams then run one after the other,
// this one first: waiting longer would only delay k_gjk_epa).  The result
// never depends on which kernel takes a run, and every wait is bounded.
__global__ __launch_bounds__(kContThreads) void k_epa_early(
    Bodies B, World W, Counters* c, const uint32_t* __restrict__ records, uint32_t record_cap,
    physgpu_narrowphase_options opt, physgpu_manifold* __restrict__ slots,
    uint32_t* __restrict__ hit, uint64_t* __restrict__ partials, uint64_t* __restrict__ unresolved,
    uint32_t* __restrict__ flags, const uint32_t* __restrict__ epa_epoch, bool solo_sphere_hull)
{
    __shared__ ContShared S;
    __shared__ uint32_t cand, take, stop;
    __shared__ unsigned long long finished;
    const uint32_t t = threadIdx.x;
    const uint32_t epoch = *epa_epoch;
    uint32_t* ready = flags;
    uint32_t* claim = flags + record_cap;
    // Work items of k_gjk_epa (its own count).
    const uint32_t c3 = c->class_count[3], c4 = c->class_count[4];
    const uint32_t total = c3 + c4 + c->class_count[5];
    const uint32_t solo_items = solo_sphere_hull ? c3 : 0;
    const uint64_t items = solo_items + (total - solo_items + kRowsPerWave - 1) / kRowsPerWave;
    unsigned long long* done = reinterpret_cast<unsigned long long*>(
        partials + size_t(kSlotEpaItemsDone) * kMaxPartialBlocks);
    const uint64_t start = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        // Thread 0 decides for the workgroup: 1 = k_epa_continue has started,
        // 2 = the wait is over.
        if (t == 0) {
            cand = kNoIndex;
            finished = 0;
            const uint64_t waited = __builtin_amdgcn_s_memrealtime() - start;
            const bool gjk_started = __hip_atomic_load(&ready[2 * record_cap + 1],
                                                       __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT) == epoch;
            stop = __hip_atomic_load(&claim[record_cap], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) == epoch
                       ? 1u
                       : (waited > kEpaEarlyPollTicks ||
                                  (!gjk_started && waited > kEpaEarlyStartTicks)
                              ? 2u
                              : 0u);
        }
        __syncthreads();
        const uint32_t stopping = stop;
        if (t < kEpaDoneSlots)
            atomicAdd(&finished, __hip_atomic_load(&done[t * kEpaDoneStride], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT));
        const uint32_t n =
            min(__hip_atomic_load(&c->epa_continued, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                record_cap);
        if (!stopping)
            for (uint32_t r = t; r < n; r += kContThreads)
                if (__hip_atomic_load(&ready[r], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) ==
                        epoch &&
                    __hip_atomic_load(&claim[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
                        epoch) {
                    atomicMin(&cand, r);
                    break;
                }
        __syncthreads();
        if (t == 0)
            take = cand != kNoIndex && atomicExch(&claim[cand], epoch) != epoch ? cand : kNoIndex;
        __syncthreads();
        const uint32_t got = take, seen = cand;
        const bool all_items = finished >= items;
        __syncthreads();  // every thread has read them before thread 0 resets them
        if (got != kNoIndex) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __builtin_amdgcn_s_setprio(3);
            continue_record(S, B, W, c, records + size_t(got) * kEpaRecordWords, record_cap, opt,
                            slots, hit, partials, unresolved);
            __builtin_amdgcn_s_setprio(0);
            continue;
        }
        if (stopping)
            break;
        if (seen != kNoIndex)
            continue;  // another workgroup claimed it first: look again
        if (all_items)
            break;
        __builtin_amdgcn_s_sleep(127);
    }
}

// ------------------------------------------------------------ hull adjoint
// Adjoint of box-hull and hull-hull manifolds through the EPA face the
// forward converged on.  One wavefront per manifold recomputes GJK and EPA
// with the forward's code (same FP32 operations, so the same polytope) and
// takes the final face (a, b, c) in its stored order: normal =
// normalize((b - a) x (c - a)), depth = normal . a, contact point = midpoint
// of the two support points along the normal.  Every polytope vertex is
// support_A - support_B of one hull point or box corner of each shape; the
// face vertices' features are found by recomputing that difference for every
// feature pair and matching it bit for bit.  With the features held fixed the
// outputs are smooth in the poses and are differentiated in reverse mode.
// Sphere-hull pairs (vertices depend on the directions of earlier faces) are
// not differentiated and are counted in `skipped`.
struct Feature {
    bool ok;
    f3 a, b;  // collider-local points of A and B
};

__device__ __forceinline__ uint32_t feature_count(const Shape& s)
{
    return s.type == PHYSGPU_HULL ? s.count : 8u;
}

__device__ __forceinline__ f3 feature_point(const Bodies& B, const Shape& s, uint32_t i)
{
    if (s.type == PHYSGPU_HULL) {
        const uint32_t p = s.first + i, nh = B.nh;
        return {B.hpts[p], B.hpts[nh + p], B.hpts[2 * nh + p]};
    }
    return {(i & 1) ? s.data.x : -s.data.x, (i & 2) ? s.data.y : -s.data.y,
            (i & 4) ? s.data.z : -s.data.z};
}

// Feature pair whose support difference is bit-identical to v (the lowest
// pair index when several are).
__device__ Feature match_vertex(const Bodies& B, const Shape& sa, const pose& ta, const Shape& sb,
                                const pose& tb, f3 v)
{
    const uint32_t na = feature_count(sa), nb = feature_count(sb), total = na * nb;
    uint32_t best = kNoIndex;
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t c = base + lane_id();
        bool eq = false;
        if (c < total) {
            const f3 w = sub(to_world(feature_point(B, sa, c / nb), ta),
                             to_world(feature_point(B, sb, c % nb), tb));
            eq = __float_as_uint(w.x) == __float_as_uint(v.x) &&
                 __float_as_uint(w.y) == __float_as_uint(v.y) &&
                 __float_as_uint(w.z) == __float_as_uint(v.z);
        }
        const uint64_t m = __ballot(eq);
        if (m && best == kNoIndex)
            best = base + uint32_t(__ffsll((unsigned long long)m)) - 1u;
    }
    Feature f{best != kNoIndex, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    if (f.ok) {
        f.a = feature_point(B, sa, best / nb);
        f.b = feature_point(B, sb, best % nb);
    }
    return f;
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
void k_hull_adjoint(Bodies B, const physgpu_manifold* __restrict__ man, const Counters* c,
                    uint32_t m, physgpu_narrowphase_options opt, const float* __restrict__ g_normal,
                    const float* __restrict__ g_depth, const float* __restrict__ g_point,
                    float* __restrict__ geom_normal, float* __restrict__ geom_depth,
                    float* __restrict__ geom_point, float* __restrict__ grad_pair,
                    float* __restrict__ grad_pose, uint32_t* __restrict__ skipped)
{
    __shared__ EpaShared shared[kEpaWavesPerBlock];
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    EpaShared& S = shared[wave];
    const uint32_t count = min(m, c->manifold_count);
    const uint32_t stride = gridDim.x * kEpaWavesPerBlock;
    uint32_t skip = 0;
    for (uint32_t t = blockIdx.x * kEpaWavesPerBlock + wave; t < count; t += stride) {
        const physgpu_manifold& mf = man[t];
        const uint32_t cls = mf.shape_pair;
        if (cls < PHYSGPU_PAIR_SPHERE_HULL)
            continue;
        if (cls == PHYSGPU_PAIR_SPHERE_HULL) {
            ++skip;
            continue;
        }
        const uint32_t a = mf.first, b = mf.second;
        const Shape sa = load_shape(B, a), sb = load_shape(B, b);
        const pose ta = load_pose(B, a), tb = load_pose(B, b);
        Simplex sx;
        if (!gjk(B, sa, ta, sb, tb, opt.gjk_max_iterations, sx)) {
            ++skip;
            continue;
        }
        const WaveHull ha = wave_hull(B, sa, ta), hb = wave_hull(B, sb, tb);
        f3 n;
        float depth;
        uint32_t face = 0;
        const int st = epa_wave(S, B, sa, ta, ha, sb, tb, hb, sx.p, opt.epa_max_iterations,
                                opt.epa_tolerance, n, depth, &face);
        __builtin_amdgcn_s_setprio(0);
        wave_sync();
        if (st != 1) {
            ++skip;
            continue;
        }
        f3 la, lb;
        const f3 w = scale(add(support_wave(B, sa, ta, n, ha, &la),
                               support_wave(B, sb, tb, neg(n), hb, &lb)),
                           0.5f);
        const f3 v0 = lds_vertex(S, face & 0xFFu), v1 = lds_vertex(S, (face >> 8) & 0xFFu),
                 v2 = lds_vertex(S, (face >> 16) & 0xFFu);
        const Feature f0 = match_vertex(B, sa, ta, sb, tb, v0);
        const Feature f1 = match_vertex(B, sa, ta, sb, tb, v1);
        const Feature f2 = match_vertex(B, sa, ta, sb, tb, v2);
        const f3 e1 = sub(v1, v0), e2 = sub(v2, v0);
        const f3 u = cross(e1, e2);
        const float len = length(u);
        if (!(f0.ok && f1.ok && f2.ok) || !(len > 1e-8f)) {
            ++skip;
            continue;
        }
        // Seeds.
        f3 gn{0.0f, 0.0f, 0.0f}, gw{0.0f, 0.0f, 0.0f};
        float gd = 0.0f;
        if (g_normal)
            gn = {g_normal[t], g_normal[m + t], g_normal[2 * m + t]};
        if (g_depth)
            gd = g_depth[t];
        if (g_point)
            gw = {g_point[t], g_point[m + t], g_point[2 * m + t]};
        PoseGrad ga{{0, 0, 0}, {0, 0, 0, 0}}, gb{{0, 0, 0}, {0, 0, 0, 0}};
        // depth = n . v0; n = u / |u|; u = e1 x e2.
        f3 gv0 = scale(n, gd);
        const f3 gu = unit_vjp(n, len, add(gn, scale(v0, gd)));
        const f3 ge1 = cross(e2, gu), ge2 = cross(gu, e1);
        const f3 gv1 = ge1, gv2 = ge2;
        gv0 = sub(gv0, add(ge1, ge2));
        const Feature fs[3] = {f0, f1, f2};
        const f3 gv[3] = {gv0, gv1, gv2};
        for (int k = 0; k < 3; ++k) {
            // v_k = (ta.p + R_A a_k) - (tb.p + R_B b_k)
            acc(ga.p, gv[k]);
            rotate_vjp(ta.q, fs[k].a, gv[k], ga.q);
            gb.p = sub(gb.p, gv[k]);
            rotate_vjp(tb.q, fs[k].b, neg(gv[k]), gb.q);
        }
        // w = (ta.p + R_A la + tb.p + R_B lb) / 2
        const f3 hw = scale(gw, 0.5f);
        acc(ga.p, hw);
        rotate_vjp(ta.q, la, hw, ga.q);
        acc(gb.p, hw);
        rotate_vjp(tb.q, lb, hw, gb.q);
        project_tangent(ga.q, ta.q);
        project_tangent(gb.q, tb.q);
        if (lane == 0) {
            if (geom_normal) {
                geom_normal[t] = n.x;
                geom_normal[m + t] = n.y;
                geom_normal[2 * m + t] = n.z;
            }
            if (geom_depth)
                geom_depth[t] = depth;
            if (geom_point) {
                geom_point[t] = w.x;
                geom_point[m + t] = w.y;
                geom_point[2 * m + t] = w.z;
            }
            const float va[7] = {ga.p.x, ga.p.y, ga.p.z, ga.q.x, ga.q.y, ga.q.z, ga.q.w};
            const float vb[7] = {gb.p.x, gb.p.y, gb.p.z, gb.q.x, gb.q.y, gb.q.z, gb.q.w};
            for (int q = 0; q < 7; ++q) {
                if (grad_pair) {
                    grad_pair[size_t(q) * m + t] = va[q];
                    grad_pair[size_t(7 + q) * m + t] = vb[q];
                }
                if (grad_pose) {
                    atomicAdd(&grad_pose[size_t(q) * B.n + a], va[q]);
                    atomicAdd(&grad_pose[size_t(q) * B.n + b], vb[q]);
                }
            }
        }
    }
    if (lane == 0 && skip)
        atomicAdd(skipped, skip);
}

// Compaction: hits per tile of kScanTile candidates.
static_assert(sizeof(physgpu_manifold) == 160 && offsetof(physgpu_manifold, pair_index) == 144,
              "k_compact copies manifolds as 10 x 16 bytes and rewrites word 36");

// Compaction in canonical order: position k of the canonical list is
// candidate perm[k] of the narrow phase's input (k itself without perm).
__device__ __forceinline__ uint32_t candidate_at(const uint32_t* __restrict__ perm, uint32_t k)
{
    return perm ? perm[k] : k;
}

// Compaction tiles are kScanTile = kBlock candidates in canonical order, one
// per thread.
static_assert(kScanTile == kBlock, "one compaction tile per workgroup pass");

__global__ __launch_bounds__(kBlock) void k_tile_hits(const uint32_t* __restrict__ hit,
                                                      const uint32_t* __restrict__ perm,
                                                      Counters* c, uint64_t cap,
                                                      uint32_t* __restrict__ tile_sums)
{
    __shared__ uint32_t part[kBlock / 64];
    // Pairs held (the broad phase's count clamped to the capacity), for this
    // and the kernels after it.
    const unsigned long long emitted = c->pair_count;
    const uint32_t count = uint32_t(emitted < cap ? emitted : cap);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        c->sort_count = count;
    const uint32_t tiles = (count + kScanTile - 1) / kScanTile;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t k = tile * kScanTile + threadIdx.x;
        const bool f = k < count && hit[candidate_at(perm, k)] != 0;
        const uint32_t s = (uint32_t)__popcll(__ballot(f));
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

// Workgroup 0: exclusive scan of the tile hit counts and the manifold count;
// workgroups 1.. fold the narrow phase's partial slots (list entry b - 1).
__global__ __launch_bounds__(kFoldThreads) void k_tile_scan(uint32_t* __restrict__ tile_sums,
                                                            Counters* c,
                                                            const uint64_t* __restrict__ partials,
                                                            FoldList fold)
{
    static_assert(kFoldThreads == 1024, "one scan entry per thread of 1024");
    if (blockIdx.x > 0) {
        fold_entry(partials, fold.e[blockIdx.x - 1], c);
        return;
    }
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

// Ordered gather: thread t of a tile owns candidate t; each wavefront lists
// its hits in LDS and copies their manifolds with all 64 lanes, one 16-byte
// word per lane (10 words per manifold).
__global__ __launch_bounds__(kBlock) void k_compact(const uint32_t* __restrict__ hit,
                                                    const uint32_t* __restrict__ perm,
                                                    const Counters* c,
                                                    const uint32_t* __restrict__ tile_sums,
                                                    const physgpu_manifold* __restrict__ slots,
                                                    physgpu_manifold* __restrict__ out)
{
    constexpr uint32_t kWords = sizeof(physgpu_manifold) / sizeof(uint4);
    static_assert(kWords == 10, "manifold = 10 x 16 bytes");
    __shared__ uint32_t wave_tot[kBlock / 64];
    __shared__ uint32_t src_of[kBlock], pos_of[kBlock];
    const uint32_t count = c->sort_count;
    const uint32_t tiles = (count + kScanTile - 1) / kScanTile;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t* my_src = src_of + wave * 64;
    uint32_t* my_pos = pos_of + wave * 64;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t k = tile * kScanTile + threadIdx.x;
        uint32_t src = 0;
        bool f = false;
        if (k < count) {
            src = candidate_at(perm, k);
            f = hit[src] != 0;
        }
        const uint64_t m = __ballot(f);
        const uint32_t hits = (uint32_t)__popcll(m);
        __syncthreads();
        if (lane == 0)
            wave_tot[wave] = hits;
        if (f) {
            const uint32_t r = (uint32_t)__popcll(m & lanes_below(lane));
            my_src[r] = src;
            my_pos[r] = k;  // pair_index: the candidate's canonical position
        }
        __syncthreads();
        uint32_t base = tile_sums[tile];
        for (uint32_t w = 0; w < wave; ++w)
            base += wave_tot[w];
        for (uint32_t i = lane; i < hits * kWords; i += 64) {
            const uint32_t j = i / kWords, q = i - kWords * j;
            uint4 v = reinterpret_cast<const uint4*>(slots + my_src[j])[q];
            // Word 36 (x of the last 16 bytes) is pair_index.
            if (q == kWords - 1)
                v.x = my_pos[j];
            reinterpret_cast<uint4*>(out + base + j)[q] = v;
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
        c->epa_continued = 0;
        c->epa_unresolved = 0;
        c->manifold_count = 0;
    }
}

}  // namespace

void launch_hull_adjoint(physgpu_context* ctx, uint32_t m, const float* g_normal,
                         const float* g_depth, const float* g_point, float* geom_normal,
                         float* geom_depth, float* geom_point, float* grad_pair, float* grad_pose,
                         uint32_t* skipped)
{
    if (!ctx->has_hulls || m == 0)
        return;
    Bodies B{ctx->n, ctx->pose4, ctx->shape4, ctx->hull_points, ctx->n_hull_points};
    Timer t;
    time_begin(ctx, "np_hull_adjoint", t);
    hipLaunchKernelGGL(k_hull_adjoint, dim3(std::min<uint32_t>(div_up(m, kEpaWavesPerBlock), 2048)),
                       dim3(kBlock), 0, ctx->stream, B, ctx->manifolds, ctx->counters, m,
                       ctx->np_options, g_normal, g_depth, g_point, geom_normal, geom_depth,
                       geom_point, grad_pair, grad_pose, skipped);
    time_end(ctx, t);
}

void launch_np_begin(physgpu_context* ctx)
{
    hipLaunchKernelGGL(k_np_begin, dim3(1), dim3(64), 0, ctx->stream, ctx->counters);
}

void launch_narrowphase(physgpu_context* ctx, bool world, bool sort_on_side)
{
    const uint64_t* pairs = ctx->np_pairs;
    Bodies B{ctx->n, ctx->pose4, ctx->shape4, ctx->hull_points, ctx->n_hull_points};
    World W{nullptr, nullptr, nullptr, nullptr};
    if (world && ctx->has_world)
        W = World{ctx->world_has_body ? ctx->body : nullptr,
                  ctx->world_has_local ? ctx->local_pose : nullptr,
                  ctx->world_has_friction ? ctx->friction : nullptr,
                  ctx->world_has_restitution ? ctx->restitution : nullptr};
    physgpu_narrowphase_options opt = ctx->np_options;
    const uint32_t grid = kStrideGrid;
    const physgpu_tuning& tu = ctx->tuning;
    Timer t;
    // The canonical pair order is built on the side stream meanwhile (from
    // the broad phase's pairs); the compaction reads the results through its
    // permutation.  The classification is enqueued first, so that a captured
    // step keeps it on the broad phase's queue, and the sort after the SS /
    // SB / BB kernels, which it would otherwise delay.
    if (sort_on_side)
        hipEventRecord(ctx->sort_event, ctx->stream);
    // A grid collision step has classified its pairs in the broad phase's
    // query (and started the EPA epoch with its key pass).
    if (!ctx->np_classified) {
        time_begin(ctx, "np_classify", t);
        hipLaunchKernelGGL(k_classify, dim3(kClassifyGrid), dim3(kClassifyThreads), 0,
                           ctx->stream, B, W, pairs, ctx->counters, ctx->buckets, ctx->cap_p,
                           ctx->hit, ctx->partials + size_t(kEpaPartialFirst) * kMaxPartialBlocks,
                           ctx->epa_epoch);
        time_end(ctx, t);
    }
    ctx->np_classified = false;
    // With hulls, GJK + EPA run here while the SS / SB / BB kernels and then
    // the canonical sort run on the side stream.
    const bool fork = ctx->has_hulls;
    hipStream_t cs = ctx->stream;
    if (fork) {
        hipEventRecord(ctx->fork_event, ctx->stream);
        hipStreamWaitEvent(ctx->side_stream, ctx->fork_event, 0);
        cs = ctx->side_stream;
        // Runs handed over past the LDS bounds are continued while GJK + EPA
        // still runs, by a consumer on a third stream, submitted first so its
        // workgroups hold CUs before GJK + EPA fills them.
        hipStreamWaitEvent(ctx->early_stream, ctx->fork_event, 0);
        time_begin(ctx, "np_epa_early", t, ctx->early_stream);
        hipLaunchKernelGGL(k_epa_early, dim3(std::min(ctx->epa_record_cap, kEpaEarlyGroups)),
                           dim3(kContThreads), 0, ctx->early_stream, B, W, ctx->counters,
                           ctx->epa_records, ctx->epa_record_cap, opt, ctx->slots, ctx->hit,
                           ctx->partials, ctx->epa_unresolved_pairs, ctx->epa_flags,
                           ctx->epa_epoch, tu.sphere_hull_per_wave == 1);
        time_end(ctx, t);
        hipEventRecord(ctx->early_join_event, ctx->early_stream);
        time_begin(ctx, "np_gjk_epa", t);
        hipLaunchKernelGGL(k_gjk_epa, dim3(tu.epa_groups), dim3(kEpaBlock), 0, ctx->stream, B,
                           W, pairs, ctx->counters, ctx->buckets, ctx->cap_p, opt, ctx->slots,
                           ctx->hit, ctx->partials, tu.sphere_hull_per_wave == 1,
                           ctx->epa_records, ctx->epa_record_cap, ctx->epa_unresolved_pairs,
                           ctx->epa_flags, ctx->epa_epoch);
        time_end(ctx, t);
        // Runs past the LDS polytope bounds that k_epa_early (side stream)
        // has not taken (none on most steps: every workgroup exits at once).
        time_begin(ctx, "np_epa_continue", t);
        hipLaunchKernelGGL(k_epa_continue, dim3(std::min(ctx->epa_record_cap, kEpaContinueGroups)),
                           dim3(kContThreads), 0, ctx->stream, B, W, ctx->counters, ctx->epa_records,
                           ctx->epa_record_cap, opt, ctx->slots, ctx->hit, ctx->partials,
                           ctx->epa_unresolved_pairs, ctx->epa_flags + ctx->epa_record_cap,
                           ctx->epa_epoch);
        time_end(ctx, t);
    }
    time_begin(ctx, "np_sphere_sphere", t, cs);
    hipLaunchKernelGGL(k_class<0>, dim3(tu.class_groups[0]), dim3(tu.class_block[0]), 0, cs, B,
                       W, pairs, ctx->counters, ctx->buckets, ctx->cap_p, opt, ctx->slots,
                       ctx->hit, ctx->partials + size_t(kSlotContacts + 0) * kMaxPartialBlocks);
    time_end(ctx, t);
    time_begin(ctx, "np_sphere_box", t, cs);
    hipLaunchKernelGGL(k_class<1>, dim3(tu.class_groups[1]), dim3(tu.class_block[1]), 0, cs, B,
                       W, pairs, ctx->counters, ctx->buckets, ctx->cap_p, opt, ctx->slots,
                       ctx->hit, ctx->partials + size_t(kSlotContacts + 1) * kMaxPartialBlocks);
    time_end(ctx, t);
    time_begin(ctx, "np_box_box", t, cs);
    hipLaunchKernelGGL(k_class<2>, dim3(tu.class_groups[2]), dim3(tu.class_block[2]), 0, cs, B,
                       W, pairs, ctx->counters, ctx->buckets, ctx->cap_p, opt, ctx->slots,
                       ctx->hit, ctx->partials + size_t(kSlotContacts + 2) * kMaxPartialBlocks);
    time_end(ctx, t);
    if (sort_on_side) {
        hipStreamWaitEvent(ctx->side_stream, ctx->sort_event, 0);
        launch_sort_pairs(ctx, ctx->side_stream);
    }

    if (fork || sort_on_side) {
        hipEventRecord(ctx->join_event, ctx->side_stream);
        hipStreamWaitEvent(ctx->stream, ctx->join_event, 0);
    }
    if (fork)
        hipStreamWaitEvent(ctx->stream, ctx->early_join_event, 0);
    // Contact and failure counts from the per-workgroup partials.
    FoldList fold{};
    for (uint32_t q = 0; q < 3; ++q)
        fold.e[q] = {kSlotContacts + q, tu.class_groups[q], kFoldSum32,
                     uint32_t(offsetof(Counters, class_contacts) + 4 * q)};
    fold.count = 3;
    if (ctx->has_hulls) {
        const uint32_t epa_slots = std::min(tu.epa_groups, kEpaPartials);
        for (uint32_t q = 3; q < kClasses; ++q)
            fold.e[q] = {kSlotContacts + q, epa_slots, kFoldSum32,
                         uint32_t(offsetof(Counters, class_contacts) + 4 * q)};
        fold.e[6] = {kSlotGjkFailures, epa_slots, kFoldSum32,
                     uint32_t(offsetof(Counters, gjk_failures))};
        fold.e[7] = {kSlotEpaOverflows, epa_slots, kFoldSum32,
                     uint32_t(offsetof(Counters, epa_overflow))};
        fold.count = 8;
    }
    time_begin(ctx, "np_compact", t);
    hipLaunchKernelGGL(k_tile_hits, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->hit,
                       ctx->np_perm, ctx->counters, ctx->cap_p, ctx->tile_sums);
    hipLaunchKernelGGL(k_tile_scan, dim3(1 + fold.count), dim3(kFoldThreads), 0, ctx->stream,
                       ctx->tile_sums, ctx->counters, ctx->partials, fold);
    hipLaunchKernelGGL(k_compact, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->hit,
                       ctx->np_perm, ctx->counters, ctx->tile_sums, ctx->slots, ctx->manifolds);
    time_end(ctx, t);
}

}  // namespace pg
