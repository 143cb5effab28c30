// C++ host adapter over the C ABI of libphysgpu.so (include/physgpu.h).
//
// It re-exposes phys-engine's collision interfaces with the same names,
// argument meaning and error behaviour, so C++ callers written against the
// reference keep their code:
//   BroadPhaseAlgorithm, BroadPhasePair, BroadPhaseStats   include/phys/collision/broadphase.hpp:8-42
//   BroadPhase::findCandidatePairs                         broadphase.hpp:52-54 (throws
//                                                          std::invalid_argument for DynamicTree,
//                                                          broadphase.cpp:200-201)
//   DynamicAabbTree createProxy / destroyProxy /
//     updateProxy / findCandidatePairs / size              include/phys/collision/dynamic_aabb_tree.hpp:25-32
//   NarrowPhase::generateContact and the typed statics     include/phys/collision/narrowphase.hpp:13-39
//   ContactPoint, ContactManifold                          include/phys/collision/manifold.hpp:9-34
//   Vec3, Quaternion, Transform, transform()               include/phys/math/{vec3,quaternion,transform}.hpp
// plus CollisionPipeline, the device-resident collision section of
// PhysicsWorld::step (physics_world.cpp:539-739).  Status codes of the C ABI
// become std::invalid_argument (PHYSGPU_INVALID_ARGUMENT) or
// std::runtime_error.  Header-only; link with -lphysgpu.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <variant>
#include <vector>

#include "../../include/physgpu.h"

namespace physgpu {

// ------------------------------------------------------------------ math
struct Vec3 {
    float x{}, y{}, z{};
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator*(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }

struct Quaternion {
    float x{}, y{}, z{}, w{1.0f};
    static constexpr Quaternion identity() { return {0.0f, 0.0f, 0.0f, 1.0f}; }
    // `axis` is unit length; the angle is the whole rotation about it.
    static Quaternion fromAxisAngle(Vec3 axis, float angle)
    {
        const float s = std::sin(angle * 0.5f);
        return {axis.x * s, axis.y * s, axis.z * s, std::cos(angle * 0.5f)};
    }
    Quaternion operator*(const Quaternion& o) const
    {
        return {w * o.x + x * o.w + y * o.z - z * o.y, w * o.y - x * o.z + y * o.w + z * o.x,
                w * o.z + x * o.y - y * o.x + z * o.w, w * o.w - x * o.x - y * o.y - z * o.z};
    }
    Quaternion conjugate() const { return {-x, -y, -z, w}; }
    // The full sandwich q (v, 0) q*, as the collision kernels evaluate it.
    Vec3 rotate(Vec3 v) const
    {
        const Quaternion r = (*this) * Quaternion{v.x, v.y, v.z, 0.0f} * conjugate();
        return {r.x, r.y, r.z};
    }
};

struct Transform {
    Vec3 position{};
    Quaternion orientation = Quaternion::identity();
    Transform() = default;
    Transform(Vec3 p, Quaternion q) : position(p), orientation(q) {}
    // Rotation about +z by `zAngle` radians.
    Transform(Vec3 p, float zAngle)
        : position(p), orientation(Quaternion::fromAxisAngle({0.0f, 0.0f, 1.0f}, zAngle))
    {
    }
};

inline Vec3 transform(Vec3 v, const Transform& t) { return t.orientation.rotate(v) + t.position; }

// ------------------------------------------------------------------ shapes
struct Aabb {
    Vec3 min{}, max{};
    bool overlaps(const Aabb& o) const
    {
        return min.x <= o.max.x && max.x >= o.min.x && min.y <= o.max.y && max.y >= o.min.y &&
               min.z <= o.max.z && max.z >= o.min.z;
    }
};
static_assert(sizeof(Aabb) == sizeof(physgpu_aabb), "Aabb must match physgpu_aabb");

struct Sphere {
    float radius{};
};
struct Box {
    Vec3 halfExtents{};
};
// Convex hull given by its collider-local points (GJK/EPA on the device).
struct ConvexHull {
    std::vector<Vec3> points;
};

struct RigidBodyHandle {
    uint32_t index = 0;
    uint32_t generation = 0;
};

struct Collider {
    RigidBodyHandle body{};
    Transform localTransform{};
    std::variant<Sphere, Box, ConvexHull> shape = Sphere{};
    Aabb bounds{};
};

// ------------------------------------------------------------------ records
enum class BroadPhaseAlgorithm { SweepAndPrune, UniformGrid, DynamicTree };

struct BroadPhasePair {
    std::size_t first;
    std::size_t second;
};
inline bool operator==(const BroadPhasePair& a, const BroadPhasePair& b)
{
    return a.first == b.first && a.second == b.second;
}

struct BroadPhaseStats {
    double recordBuildMs = 0.0;
    double recordSortMs = 0.0;
    double sweepMs = 0.0;
    double pairSortMs = 0.0;
    std::size_t xWindowComparisons = 0;
    std::size_t aabbPairs = 0;
    std::size_t gridComparisons = 0;
    std::size_t gridEntries = 0;
    std::size_t gridOverflowAabbs = 0;
    double gridCellSize = 0.0;
    std::size_t treeInsertions = 0;
    std::size_t treeRemovals = 0;
    std::size_t treeReinsertions = 0;
    std::size_t treeNodePairVisits = 0;
    std::size_t treeLeafChecks = 0;
    std::size_t treeHeight = 0;
    std::size_t treeProxyCount = 0;
};

struct ContactPoint {
    Vec3 localAnchorA{};
    Vec3 localAnchorB{};
    float penetration = 0.0f;
    float normalImpulse = 0.0f;
    float tangentImpulse1 = 0.0f;
    float tangentImpulse2 = 0.0f;
};

struct ContactManifold {
    RigidBodyHandle bodyA{};
    RigidBodyHandle bodyB{};
    Vec3 normal{};  // from body A toward body B
    std::array<ContactPoint, 4> points{};
    uint32_t pointCount = 0;
    float friction = 0.0f;
    float restitution = 0.0f;
};

struct NarrowPhaseStats {
    std::size_t sphereSphereCandidates = 0;
    std::size_t sphereBoxCandidates = 0;
    std::size_t boxBoxCandidates = 0;
    std::size_t sphereSphereContacts = 0;
    std::size_t sphereBoxContacts = 0;
    std::size_t boxBoxContacts = 0;
    std::size_t hullCandidates = 0;
    std::size_t hullContacts = 0;
};

// ------------------------------------------------------------------ errors
inline void check(physgpu_status st)
{
    if (st == PHYSGPU_OK)
        return;
    const std::string msg = physgpu_last_error();
    if (st == PHYSGPU_INVALID_ARGUMENT)
        throw std::invalid_argument(msg);
    throw std::runtime_error("physgpu: " + msg);
}

inline bool finiteOrdered(const Aabb& b)
{
    const float lo[3] = {b.min.x, b.min.y, b.min.z}, hi[3] = {b.max.x, b.max.y, b.max.z};
    for (int a = 0; a < 3; ++a)
        if (!(std::isfinite(lo[a]) && std::isfinite(hi[a]) && lo[a] <= hi[a]))
            return false;
    return true;
}

// ------------------------------------------------------------------ context
// One device context (device 0, its own stream), grown to the sizes asked.
class Context {
public:
    Context() = default;
    Context(uint32_t colliders, uint64_t pairs, uint32_t hullPoints) { reserve(colliders, pairs, hullPoints); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    ~Context() { physgpu_destroy(ctx_); }

    void reserve(uint32_t colliders, uint64_t pairs, uint32_t hullPoints = 0)
    {
        colliders = std::max<uint32_t>(colliders, 1);
        pairs = std::max<uint64_t>(pairs, 1024);
        if (ctx_ && colliders <= colliders_ && pairs <= pairs_ && hullPoints <= hull_)
            return;
        physgpu_destroy(ctx_);
        ctx_ = nullptr;
        colliders_ = std::max(colliders, colliders_);
        pairs_ = std::max(pairs, pairs_);
        hull_ = std::max(hullPoints, hull_);
        check(physgpu_create(&ctx_, 0, colliders_, pairs_, hull_));
    }
    physgpu_context* get() const { return ctx_; }
    uint64_t pairCapacity() const { return pairs_; }

    // A context per calling thread for the stateless calls (the reference's
    // statics are safe to call from several threads at once).
    static Context& threadDefault()
    {
        thread_local Context c;
        return c;
    }

private:
    physgpu_context* ctx_ = nullptr;
    uint32_t colliders_ = 0;
    uint64_t pairs_ = 0;
    uint32_t hull_ = 0;
};

namespace detail {

inline BroadPhaseStats convert(const physgpu_broadphase_stats& s)
{
    BroadPhaseStats o;
    o.recordBuildMs = s.record_build_ms;
    o.recordSortMs = s.record_sort_ms;
    o.sweepMs = s.sweep_ms;
    o.pairSortMs = s.pair_sort_ms;
    o.xWindowComparisons = s.x_window_comparisons;
    o.aabbPairs = s.aabb_pairs;
    o.gridComparisons = s.grid_comparisons;
    o.gridEntries = s.grid_entries;
    o.gridOverflowAabbs = s.grid_overflow_aabbs;
    o.gridCellSize = s.grid_cell_size;
    o.treeInsertions = s.tree_insertions;
    o.treeRemovals = s.tree_removals;
    o.treeReinsertions = s.tree_reinsertions;
    o.treeNodePairVisits = s.tree_node_pair_visits;
    o.treeLeafChecks = s.tree_leaf_checks;
    o.treeHeight = s.tree_height;
    o.treeProxyCount = s.tree_proxy_count;
    return o;
}

// Any algorithm, including the tree (an LBVH built for the call).
inline std::vector<BroadPhasePair> pairsOf(Context& c, const std::vector<Aabb>& bounds,
                                           BroadPhaseAlgorithm algorithm, BroadPhaseStats* stats)
{
    if (bounds.size() > std::numeric_limits<uint32_t>::max() / 2)
        throw std::invalid_argument("too many bounds");
    for (const Aabb& b : bounds)
        if (!finiteOrdered(b))
            throw std::invalid_argument("Broadphase bounds must be finite and ordered");
    const uint32_t n = uint32_t(bounds.size());
    if (n < 2) {
        if (stats)
            *stats = BroadPhaseStats{};
        return {};
    }
    c.reserve(n, std::max<uint64_t>(4ull * n, 1024));
    physgpu_broadphase_stats st{};
    for (;;) {
        std::vector<physgpu_pair> out(c.pairCapacity());
        uint64_t count = 0;
        const physgpu_status s = physgpu_find_candidate_pairs(
            c.get(), reinterpret_cast<const physgpu_aabb*>(bounds.data()), n, int(algorithm), 0.0,
            out.data(), out.size(), &count, &st);
        if (s == PHYSGPU_CAPACITY_EXCEEDED && count > c.pairCapacity()) {
            c.reserve(n, count);  // grow and run again
            continue;
        }
        check(s);
        std::vector<BroadPhasePair> pairs(count);
        for (uint64_t k = 0; k < count; ++k)
            pairs[k] = {out[k].first, out[k].second};
        if (stats)
            *stats = convert(st);
        return pairs;
    }
}

// SoA collider planes for physgpu_set_colliders.
struct ColliderPlanes {
    uint32_t n = 0;
    std::vector<float> pose, shape;
    std::vector<uint8_t> type;
    std::vector<uint32_t> hullRange;
    std::vector<float> hullPoints;
    bool hulls = false;

    void build(const std::vector<const Collider*>& cs, const std::vector<Transform>& world)
    {
        n = uint32_t(cs.size());
        pose.assign(7 * size_t(n), 0.0f);
        shape.assign(3 * size_t(n), 0.0f);
        type.assign(n, 0);
        hullRange.assign(2 * size_t(n), 0);
        std::vector<Vec3> pts;
        hulls = false;
        for (uint32_t i = 0; i < n; ++i) {
            const Transform& t = world[i];
            const float p[7] = {t.position.x, t.position.y, t.position.z, t.orientation.x,
                                t.orientation.y, t.orientation.z, t.orientation.w};
            for (int q = 0; q < 7; ++q)
                pose[q * size_t(n) + i] = p[q];
            const auto& s = cs[i]->shape;
            if (const Sphere* sp = std::get_if<Sphere>(&s)) {
                type[i] = PHYSGPU_SPHERE;
                shape[i] = sp->radius;
            } else if (const Box* bx = std::get_if<Box>(&s)) {
                type[i] = PHYSGPU_BOX;
                shape[i] = bx->halfExtents.x;
                shape[n + i] = bx->halfExtents.y;
                shape[2 * size_t(n) + i] = bx->halfExtents.z;
            } else {
                const ConvexHull& h = std::get<ConvexHull>(s);
                if (h.points.empty())
                    throw std::invalid_argument("convex hull without points");
                type[i] = PHYSGPU_HULL;
                hullRange[i] = uint32_t(pts.size());
                hullRange[n + i] = uint32_t(h.points.size());
                pts.insert(pts.end(), h.points.begin(), h.points.end());
                hulls = true;
            }
        }
        const size_t H = pts.size();
        hullPoints.assign(3 * std::max<size_t>(H, 1), 0.0f);
        for (size_t k = 0; k < H; ++k) {
            hullPoints[k] = pts[k].x;
            hullPoints[H + k] = pts[k].y;
            hullPoints[2 * H + k] = pts[k].z;
        }
    }
    uint32_t hullCount() const { return hulls ? uint32_t(hullPoints.size() / 3) : 0; }
    physgpu_colliders view() const
    {
        return {n, pose.data(), type.data(), shape.data(), hulls ? hullRange.data() : nullptr,
                hullCount(), hulls ? hullPoints.data() : nullptr};
    }
};

inline ContactManifold toManifold(const physgpu_manifold& m, RigidBodyHandle a,
                                  RigidBodyHandle b)
{
    ContactManifold cm;
    cm.bodyA = a;
    cm.bodyB = b;
    cm.normal = {m.normal[0], m.normal[1], m.normal[2]};
    cm.pointCount = m.point_count;
    for (uint32_t k = 0; k < m.point_count && k < 4; ++k) {
        cm.points[k].localAnchorA = {m.anchor_a[k][0], m.anchor_a[k][1], m.anchor_a[k][2]};
        cm.points[k].localAnchorB = {m.anchor_b[k][0], m.anchor_b[k][1], m.anchor_b[k][2]};
        cm.points[k].penetration = m.penetration[k];
    }
    cm.friction = m.friction;
    cm.restitution = m.restitution;
    return cm;
}

}  // namespace detail

// ------------------------------------------------------------------ broad phase
class BroadPhase {
public:
    // Every touching or overlapping pair of input indices, first < second, in
    // lexicographic order; `stats` is overwritten.  DynamicTree needs a
    // persistent DynamicAabbTree, as in the reference.
    static std::vector<BroadPhasePair> findCandidatePairs(
        const std::vector<Aabb>& bounds, BroadPhaseStats* stats = nullptr,
        BroadPhaseAlgorithm algorithm = BroadPhaseAlgorithm::SweepAndPrune)
    {
        if (algorithm == BroadPhaseAlgorithm::DynamicTree)
            throw std::invalid_argument("DynamicTree requires a persistent DynamicAabbTree");
        return detail::pairsOf(Context::threadDefault(), bounds, algorithm, stats);
    }
};

// Persistent proxies with the reference's fat-bounds bookkeeping
// (dynamic_aabb_tree.cpp:44-57, updateProxy at :69-81); the pair query runs
// on the device over the proxies' tight bounds (an LBVH built per query), so
// the pair set is the reference tree's: sorted, unique (min, max) user-index
// pairs of overlapping tight bounds.
class DynamicAabbTree {
public:
    using ProxyId = std::size_t;
    static constexpr ProxyId noProxy = std::numeric_limits<ProxyId>::max();

    ProxyId createProxy(const Aabb& bounds, std::size_t userIndex)
    {
        if (!finiteOrdered(bounds))
            throw std::invalid_argument("Dynamic AABB tree requires finite, ordered bounds");
        ProxyId id;
        if (!free_.empty()) {
            id = free_.back();
            free_.pop_back();
        } else {
            id = slots_.size();
            slots_.emplace_back();
        }
        slots_[id] = {true, bounds, expanded(bounds, 1.0), userIndex};
        ++count_;
        return id;
    }

    void destroyProxy(ProxyId proxy)
    {
        live(proxy);
        slots_[proxy].live = false;
        free_.push_back(proxy);
        --count_;
    }

    // True when the proxy was reinserted: its new bounds left the fat bounds,
    // or the fat bounds became more than 4x padding too large.
    bool updateProxy(ProxyId proxy, const Aabb& bounds)
    {
        live(proxy);
        if (!finiteOrdered(bounds))
            throw std::invalid_argument("Dynamic AABB tree requires finite, ordered bounds");
        Slot& s = slots_[proxy];
        s.tight = bounds;
        if (contains(s.fat, bounds) && contains(expanded(bounds, 4.0), s.fat))
            return false;
        s.fat = expanded(bounds, 1.0);
        return true;
    }

    std::vector<BroadPhasePair> findCandidatePairs(BroadPhaseStats* stats = nullptr) const
    {
        std::vector<Aabb> bounds;
        std::vector<std::size_t> users;
        for (const Slot& s : slots_)
            if (s.live) {
                bounds.push_back(s.tight);
                users.push_back(s.user);
            }
        BroadPhaseStats st;
        std::vector<BroadPhasePair> pairs;
        if (bounds.size() >= 2)
            pairs = detail::pairsOf(Context::threadDefault(), bounds,
                                    BroadPhaseAlgorithm::DynamicTree, &st);
        for (BroadPhasePair& p : pairs) {
            const std::size_t a = users[p.first], b = users[p.second];
            p = {std::min(a, b), std::max(a, b)};
        }
        std::sort(pairs.begin(), pairs.end(), [](const BroadPhasePair& l, const BroadPhasePair& r) {
            return l.first != r.first ? l.first < r.first : l.second < r.second;
        });
        pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
        st.aabbPairs = pairs.size();
        st.treeProxyCount = count_;
        if (stats)
            *stats = st;
        return pairs;
    }

    std::size_t size() const { return count_; }

private:
    struct Slot {
        bool live = false;
        Aabb tight{}, fat{};
        std::size_t user = 0;
    };
    void live(ProxyId proxy) const
    {
        if (proxy >= slots_.size() || !slots_[proxy].live)
            throw std::invalid_argument("Invalid dynamic AABB tree proxy");
    }
    static bool contains(const Aabb& outer, const Aabb& inner)
    {
        return outer.min.x <= inner.min.x && outer.min.y <= inner.min.y &&
               outer.min.z <= inner.min.z && outer.max.x >= inner.max.x &&
               outer.max.y >= inner.max.y && outer.max.z >= inner.max.z;
    }
    // Padding max(0.01, 10% of the side) x factor per axis, in double, clamped
    // to the float range.
    static Aabb expanded(const Aabb& b, double factor)
    {
        auto axis = [factor](float lo, float hi, float& outLo, float& outHi) {
            const double pad = std::max(0.01, (double(hi) - lo) * 0.1) * factor;
            const double limit = std::numeric_limits<float>::max();
            outLo = float(std::max(-limit, double(lo) - pad));
            outHi = float(std::min(limit, double(hi) + pad));
        };
        Aabb o;
        axis(b.min.x, b.max.x, o.min.x, o.max.x);
        axis(b.min.y, b.max.y, o.min.y, o.max.y);
        axis(b.min.z, b.max.z, o.min.z, o.max.z);
        return o;
    }
    std::vector<Slot> slots_;
    std::vector<ProxyId> free_;
    std::size_t count_ = 0;
};

// ------------------------------------------------------------------ narrow phase
class NarrowPhase {
public:
    // Contact between two colliders at world transforms; the manifold is reset
    // first; true on contact (narrowphase.cpp:326-374).
    static bool generateContact(const Collider& a, const Transform& ta, const Collider& b,
                                const Transform& tb, ContactManifold& manifold)
    {
        std::vector<ContactManifold> out;
        std::vector<uint8_t> hit;
        generateContacts({&a, &b}, {ta, tb}, {{0, 1}}, out, hit);
        manifold = out[0];
        return hit[0] != 0;
    }
    static bool intersectSphereSphere(const Sphere& a, const Transform& ta, const Sphere& b,
                                      const Transform& tb, ContactManifold& m)
    {
        Collider ca, cb;
        ca.shape = a;
        cb.shape = b;
        return generateContact(ca, ta, cb, tb, m);
    }
    static bool intersectSphereBox(const Sphere& s, const Transform& ts, const Box& b,
                                   const Transform& tb, ContactManifold& m)
    {
        Collider ca, cb;
        ca.shape = s;
        cb.shape = b;
        return generateContact(ca, ts, cb, tb, m);
    }
    static bool intersectBoxBox(const Box& a, const Transform& ta, const Box& b,
                                const Transform& tb, ContactManifold& m)
    {
        Collider ca, cb;
        ca.shape = a;
        cb.shape = b;
        return generateContact(ca, ta, cb, tb, m);
    }

    // Batched form: one manifold slot per pair (reset when no contact).
    static void generateContacts(const std::vector<const Collider*>& colliders,
                                 const std::vector<Transform>& world,
                                 const std::vector<BroadPhasePair>& pairs,
                                 std::vector<ContactManifold>& out, std::vector<uint8_t>& hit)
    {
        if (colliders.size() != world.size())
            throw std::invalid_argument("one world transform per collider");
        detail::ColliderPlanes planes;
        planes.build(colliders, world);
        Context& c = Context::threadDefault();
        c.reserve(planes.n, pairs.size(), planes.hullCount());
        const physgpu_colliders v = planes.view();
        check(physgpu_set_colliders(c.get(), &v, PHYSGPU_HOST));
        std::vector<physgpu_pair> pp(pairs.size());
        for (size_t k = 0; k < pairs.size(); ++k) {
            if (pairs[k].first >= planes.n || pairs[k].second >= planes.n)
                throw std::invalid_argument("pair index out of range");
            pp[k] = {uint32_t(pairs[k].first), uint32_t(pairs[k].second)};
        }
        std::vector<physgpu_manifold> m(std::max<size_t>(pairs.size(), 1));
        uint64_t count = 0;
        check(physgpu_generate_contacts(c.get(), pp.data(), pp.size(), nullptr, m.data(), m.size(),
                                        &count, nullptr));
        out.assign(pairs.size(), ContactManifold{});
        hit.assign(pairs.size(), 0);
        for (uint64_t i = 0; i < count; ++i) {
            const uint32_t k = m[i].pair_index;
            out[k] = detail::toManifold(m[i], colliders[pairs[k].first]->body,
                                        colliders[pairs[k].second]->body);
            hit[k] = 1;
        }
    }
};

// ------------------------------------------------------------------ world step
// The collision section of PhysicsWorld::step on the device: collider world
// transforms = body pose x local transform, AABB refresh, broad phase, pairs
// of one body dropped, narrow phase, body-local anchors, combined materials,
// manifolds in candidate order.
class CollisionPipeline {
public:
    CollisionPipeline(uint32_t maxColliders, uint64_t maxPairs, uint32_t maxHullPoints = 0)
        : ctx_(maxColliders, maxPairs, maxHullPoints)
    {
    }

    // Colliders with their bodies' poses and materials (per collider).
    void setColliders(const std::vector<Collider>& colliders,
                      const std::vector<Transform>& bodyPose, const std::vector<float>& friction,
                      const std::vector<float>& restitution)
    {
        const size_t n = colliders.size();
        if (bodyPose.size() != n || friction.size() != n || restitution.size() != n)
            throw std::invalid_argument("one body pose and material per collider");
        std::vector<const Collider*> ptrs(n);
        std::vector<Transform> local(n);
        for (size_t i = 0; i < n; ++i) {
            ptrs[i] = &colliders[i];
            local[i] = colliders[i].localTransform;
        }
        planes_.build(ptrs, local);  // shapes; poses are replaced by set_world
        bodies_.resize(n);
        std::vector<uint32_t> body(n);
        std::vector<float> bodyPlanes(7 * n), localPlanes(7 * n);
        for (size_t i = 0; i < n; ++i) {
            bodies_[i] = colliders[i].body;
            body[i] = colliders[i].body.index;
            const Transform& b = bodyPose[i];
            const Transform& l = colliders[i].localTransform;
            const float bp[7] = {b.position.x, b.position.y, b.position.z, b.orientation.x,
                                 b.orientation.y, b.orientation.z, b.orientation.w};
            const float lp[7] = {l.position.x, l.position.y, l.position.z, l.orientation.x,
                                 l.orientation.y, l.orientation.z, l.orientation.w};
            for (int q = 0; q < 7; ++q) {
                bodyPlanes[q * n + i] = bp[q];
                localPlanes[q * n + i] = lp[q];
            }
        }
        const physgpu_colliders v = planes_.view();
        check(physgpu_set_colliders(ctx_.get(), &v, PHYSGPU_HOST));
        const physgpu_world_extras ex{body.data(), localPlanes.data(), friction.data(),
                                      restitution.data()};
        check(physgpu_set_world(ctx_.get(), bodyPlanes.data(), &ex, PHYSGPU_HOST));
    }

    std::vector<ContactManifold> collide(BroadPhaseAlgorithm algorithm,
                                         NarrowPhaseStats* stats = nullptr)
    {
        uint64_t pairs = 0, count = 0;
        check(physgpu_collide(ctx_.get(), int(algorithm), 0.0, nullptr, &pairs, &count));
        std::vector<physgpu_manifold> m(std::max<uint64_t>(count, 1));
        check(physgpu_download_manifolds(ctx_.get(), m.data(), count));
        std::vector<ContactManifold> out(count);
        for (uint64_t i = 0; i < count; ++i)
            out[i] = detail::toManifold(m[i], bodies_[m[i].first], bodies_[m[i].second]);
        if (stats) {
            physgpu_narrowphase_stats s{};
            check(physgpu_narrowphase_stats_get(ctx_.get(), &s));
            *stats = {s.sphere_sphere_candidates, s.sphere_box_candidates, s.box_box_candidates,
                      s.sphere_sphere_contacts,   s.sphere_box_contacts,   s.box_box_contacts,
                      s.hull_candidates,          s.hull_contacts};
        }
        lastPairs_ = pairs;
        return out;
    }
    uint64_t lastCandidatePairs() const { return lastPairs_; }
    physgpu_context* context() const { return ctx_.get(); }

private:
    Context ctx_;
    detail::ColliderPlanes planes_;
    std::vector<RigidBodyHandle> bodies_;
    uint64_t lastPairs_ = 0;
};

}  // namespace physgpu
