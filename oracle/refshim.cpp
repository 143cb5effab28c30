// Test infrastructure only: a C ABI over the reference engine built from
// /root/reference sources by oracle/Makefile into oracle/_ref/libphysref.so.
// Only tests/, the golden-fixture generator and bench.py's CPU-baseline leg
// load it.  Nothing here is shipped or measured as the product.
//
// Every entry point calls the reference's own functions:
//   Aabb::fromCollider               src/collision/aabb.cpp:8-32
//   BroadPhase::findCandidatePairs   src/collision/broadphase.cpp:186-193
//   detail::findCandidatePairs       src/collision/broadphase.cpp:195-291 (threaded SAP)
//   DynamicAabbTree                  src/collision/dynamic_aabb_tree.cpp:108-336
//   NarrowPhase::generateContact     src/collision/narrowphase.cpp:326-374
//   intersectGjk / computeEpa        src/collision/gjk.cpp:128-158, epa.cpp:64-130
//   PhysicsWorld::step               src/world/physics_world.cpp:518-782
// The GJK/EPA support maps are defined here because the reference has no
// caller for GJK/EPA and no hull shape (convex_hull.hpp:5 is a forward
// declaration); they are the same support maps the product uses.
#include <phys/collision/aabb.hpp>
#include <phys/collision/broadphase.hpp>
#include <phys/collision/collider.hpp>
#include <phys/collision/dynamic_aabb_tree.hpp>
#include <phys/collision/epa.hpp>
#include <phys/collision/gjk.hpp>
#include <phys/collision/narrowphase.hpp>
#include <phys/math/math_utils.hpp>
#include <phys/world/physics_world.hpp>

#include "world/step_workspace.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <vector>

#include "../include/physgpu.h"

using namespace phys;

namespace {

struct ColliderView {
    const physgpu_colliders* c;
    Transform transform(uint32_t i) const {
        const uint32_t n = c->count;
        Transform t;
        t.position = {c->pose[0 * n + i], c->pose[1 * n + i], c->pose[2 * n + i]};
        t.orientation = {c->pose[3 * n + i], c->pose[4 * n + i], c->pose[5 * n + i],
                         c->pose[6 * n + i]};
        return t;
    }
    int type(uint32_t i) const { return c->shape_type[i]; }
    Vec3 data(uint32_t i) const {
        const uint32_t n = c->count;
        return {c->shape_data[i], c->shape_data[n + i], c->shape_data[2 * n + i]};
    }
    Collider collider(uint32_t i) const {
        Collider col{};
        col.body = {i, 0};
        if (type(i) == PHYSGPU_SPHERE)
            col.shape = Sphere{data(i).x};
        else
            col.shape = Box{data(i)};
        return col;
    }
};

thread_local std::string lastError;

int fail(const std::exception& e)
{
    lastError = e.what();
    return PHYSGPU_INVALID_ARGUMENT;
}

void fillBroadStats(const BroadPhaseStats& s, physgpu_broadphase_stats* out)
{
    if (!out)
        return;
    out->record_build_ms = s.recordBuildMs;
    out->record_sort_ms = s.recordSortMs;
    out->sweep_ms = s.sweepMs;
    out->pair_sort_ms = s.pairSortMs;
    out->x_window_comparisons = s.xWindowComparisons;
    out->aabb_pairs = s.aabbPairs;
    out->grid_comparisons = s.gridComparisons;
    out->grid_entries = s.gridEntries;
    out->grid_overflow_aabbs = s.gridOverflowAabbs;
    out->grid_cell_size = s.gridCellSize;
    out->tree_insertions = s.treeInsertions;
    out->tree_removals = s.treeRemovals;
    out->tree_reinsertions = s.treeReinsertions;
    out->tree_node_pair_visits = s.treeNodePairVisits;
    out->tree_leaf_checks = s.treeLeafChecks;
    out->tree_height = s.treeHeight;
    out->tree_proxy_count = s.treeProxyCount;
}

void writeManifold(const ContactManifold& m, physgpu_manifold* out)
{
    std::memset(out, 0, sizeof(*out));
    out->first = m.bodyA.index;
    out->second = m.bodyB.index;
    out->normal[0] = m.normal.x;
    out->normal[1] = m.normal.y;
    out->normal[2] = m.normal.z;
    out->point_count = m.pointCount;
    for (uint32_t k = 0; k < 4; ++k) {
        const ContactPoint& p = m.points[k];
        out->anchor_a[k][0] = p.localAnchorA.x;
        out->anchor_a[k][1] = p.localAnchorA.y;
        out->anchor_a[k][2] = p.localAnchorA.z;
        out->anchor_b[k][0] = p.localAnchorB.x;
        out->anchor_b[k][1] = p.localAnchorB.y;
        out->anchor_b[k][2] = p.localAnchorB.z;
        out->penetration[k] = p.penetration;
    }
    out->friction = m.friction;
    out->restitution = m.restitution;
}

// Support maps for GJK/EPA (builder-defined; the reference supplies none).
Vec3 rotateInverse(const Quaternion& q, Vec3 v) { return q.conjugate().rotate(v); }

SupportFunction supportFor(const ColliderView& v, uint32_t i)
{
    Transform t = v.transform(i);
    int type = v.type(i);
    if (type == PHYSGPU_SPHERE) {
        float r = v.data(i).x;
        return [t, r](const Vec3& d) { return t.position + Math3d::normalize(d) * r; };
    }
    if (type == PHYSGPU_BOX) {
        Vec3 h = v.data(i);
        return [t, h](const Vec3& d) {
            Vec3 ld = rotateInverse(t.orientation, d);
            Vec3 lp{ld.x >= 0.0f ? h.x : -h.x, ld.y >= 0.0f ? h.y : -h.y,
                    ld.z >= 0.0f ? h.z : -h.z};
            return t.position + t.orientation.rotate(lp);
        };
    }
    const uint32_t n = v.c->count;
    uint32_t first = v.c->hull_range[i];
    uint32_t count = v.c->hull_range[n + i];
    std::vector<Vec3> pts(count);
    const uint32_t H = v.c->hull_point_count;
    for (uint32_t k = 0; k < count; ++k)
        pts[k] = {v.c->hull_points[first + k], v.c->hull_points[H + first + k],
                  v.c->hull_points[2 * H + first + k]};
    return [t, pts](const Vec3& d) {
        Vec3 ld = rotateInverse(t.orientation, d);
        uint32_t best = 0;
        float bestDot = Math3d::dot(pts[0], ld);
        for (uint32_t k = 1; k < pts.size(); ++k) {
            float s = Math3d::dot(pts[k], ld);
            if (s > bestDot) {
                bestDot = s;
                best = k;
            }
        }
        return t.position + t.orientation.rotate(pts[best]);
    };
}

} // namespace

extern "C" {

const char* ref_last_error() { return lastError.c_str(); }

int ref_compute_aabbs(const physgpu_colliders* c, physgpu_aabb* out)
{
    try {
        ColliderView v{c};
        for (uint32_t i = 0; i < c->count; ++i) {
            if (v.type(i) == PHYSGPU_HULL)
                throw std::invalid_argument("the reference has no hull shape");
            Aabb b = Aabb::fromCollider(v.collider(i), v.transform(i));
            out[i] = {{b.min.x, b.min.y, b.min.z}, {b.max.x, b.max.y, b.max.z}};
        }
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// algorithm: 0 SAP, 1 grid, 2 DynamicAabbTree (proxies created in index order).
// workers > 1 runs the world's threaded SAP path (detail::findCandidatePairs
// with a ParallelFor pool), as PhysicsWorld::step does.
int ref_find_candidate_pairs(const physgpu_aabb* bounds, uint32_t n, int algorithm,
                             uint32_t workers, physgpu_pair* out, uint64_t capacity,
                             uint64_t* count, physgpu_broadphase_stats* statsOut)
{
    try {
        std::vector<Aabb> b(n);
        for (uint32_t i = 0; i < n; ++i)
            b[i] = {{bounds[i].min[0], bounds[i].min[1], bounds[i].min[2]},
                    {bounds[i].max[0], bounds[i].max[1], bounds[i].max[2]}};
        BroadPhaseStats stats{};
        std::vector<BroadPhasePair> pairs;
        if (algorithm == 2) {
            DynamicAabbTree tree;
            for (uint32_t i = 0; i < n; ++i)
                tree.createProxy(b[i], i);
            pairs = tree.findCandidatePairs(&stats);
        } else if (workers > 1) {
            detail::BroadPhaseWorkspace ws;
            detail::ParallelFor pool;
            pool.configure(workers);
            detail::findCandidatePairs(b, pairs, &stats,
                                       static_cast<BroadPhaseAlgorithm>(algorithm), ws, &pool,
                                       workers);
        } else {
            pairs = BroadPhase::findCandidatePairs(b, &stats,
                                                   static_cast<BroadPhaseAlgorithm>(algorithm));
        }
        *count = pairs.size();
        if (out)
            for (size_t k = 0; k < pairs.size() && k < capacity; ++k)
                out[k] = {static_cast<uint32_t>(pairs[k].first),
                          static_cast<uint32_t>(pairs[k].second)};
        fillBroadStats(stats, statsOut);
        return pairs.size() > capacity ? PHYSGPU_CAPACITY_EXCEEDED : PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// NarrowPhase::generateContact per pair; one output slot per pair (not
// compacted), hit[k] = return value.
int ref_generate_contacts(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                          physgpu_manifold* out, uint8_t* hit)
{
    try {
        ColliderView v{c};
        for (uint64_t k = 0; k < n; ++k) {
            uint32_t a = pairs[k].first, b = pairs[k].second;
            ContactManifold m;
            bool h = NarrowPhase::generateContact(v.collider(a), v.transform(a), v.collider(b),
                                                  v.transform(b), m);
            hit[k] = h ? 1 : 0;
            writeManifold(m, &out[k]);
            out[k].pair_index = static_cast<uint32_t>(k);
            out[k].shape_pair = static_cast<uint32_t>(v.type(a) + v.type(b));
        }
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// intersectGjk then computeEpa with the support maps above.
// status[k]: 0 = GJK reports no intersection, 1 = EPA converged,
// 2 = GJK intersection but EPA failed.
int ref_gjk_epa(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                uint32_t gjkIterations, uint32_t epaIterations, float tolerance,
                int32_t* status, float* normal, float* depth, float* simplex)
{
    try {
        ColliderView v{c};
        for (uint64_t k = 0; k < n; ++k) {
            SupportFunction sa = supportFor(v, pairs[k].first);
            SupportFunction sb = supportFor(v, pairs[k].second);
            GjkSimplex s;
            bool hit = intersectGjk(sa, sb, &s, gjkIterations);
            if (simplex) {
                for (int p = 0; p < 4; ++p) {
                    simplex[k * 13 + p * 3 + 0] = s.points[p].x;
                    simplex[k * 13 + p * 3 + 1] = s.points[p].y;
                    simplex[k * 13 + p * 3 + 2] = s.points[p].z;
                }
                simplex[k * 13 + 12] = static_cast<float>(s.size);
            }
            normal[3 * k] = normal[3 * k + 1] = normal[3 * k + 2] = 0.0f;
            depth[k] = 0.0f;
            if (!hit) {
                status[k] = 0;
                continue;
            }
            EpaResult r;
            bool ok = computeEpa(sa, sb, s, r, epaIterations, tolerance);
            status[k] = ok ? 1 : 2;
            normal[3 * k] = r.normal.x;
            normal[3 * k + 1] = r.normal.y;
            normal[3 * k + 2] = r.normal.z;
            depth[k] = r.depth;
        }
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// World-level collision section.  One static body per collider at
// body_position (the reference has no rotation setter, so body rotations are
// identity); the collider carries local_pose (7 SoA planes, may be NULL) and
// the shape of `c` (whose pose is ignored here).  Runs warmup + steps calls of
// PhysicsWorld::step and returns the contacts of the last one.
// timing[0..5] = median broadPhaseMs, narrowPhaseMs, aabb-refresh ms (total
// minus the other stages), totalMs, candidatePairs, possiblePairs.
// npStats = 6 NarrowPhaseStats counters.
static int worldCollide(const physgpu_colliders* c, const float* bodyPosition,
                        const uint32_t* colliderBody,
                      const float* localPose, const float* friction, const float* restitution,
                      int algorithm, uint32_t bpWorkers, uint32_t npWorkers, uint32_t warmup,
                      uint32_t steps, physgpu_manifold* out, uint64_t capacity, uint64_t* count,
                      double* timing, uint64_t* npStats)
{
    try {
        ColliderView v{c};
        const uint32_t n = c->count;
        PhysicsWorld world;
        world.gravity = {0.0f, 0.0f, 0.0f};
        world.broadPhaseAlgorithm = static_cast<BroadPhaseAlgorithm>(algorithm);
        world.setBroadPhaseWorkerCount(bpWorkers);
        world.setNarrowPhaseWorkerCount(npWorkers);
        // One body per collider, or per distinct colliderBody value (added in
        // the order of each body's first collider, with that collider's
        // position and materials).
        std::vector<RigidBodyHandle> handle(n);
        std::vector<int64_t> made(n, -1);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t key = colliderBody ? colliderBody[i] : i;
            if (made[key] < 0) {
                RigidBody body({bodyPosition[i], bodyPosition[n + i], bodyPosition[2 * n + i]},
                               1.0f, restitution ? restitution[i] : 0.0f,
                               friction ? friction[i] : 0.0f, true);
                made[key] = int64_t(i);
                handle[key] = world.addBody(body);
            }
        }
        for (uint32_t i = 0; i < n; ++i) {
            Collider col = v.collider(i);
            col.body = handle[colliderBody ? colliderBody[i] : i];
            if (localPose) {
                col.localTransform.position = {localPose[i], localPose[n + i],
                                               localPose[2 * n + i]};
                col.localTransform.orientation = {localPose[3 * n + i], localPose[4 * n + i],
                                                  localPose[5 * n + i], localPose[6 * n + i]};
            }
            world.addCollider(col);
        }
        std::vector<double> bp, np, aabb, total;
        for (uint32_t s = 0; s < warmup + steps; ++s) {
            world.step(1.0f / 120.0f);
            if (s < warmup)
                continue;
            const StepStats& st = world.lastStepStats();
            bp.push_back(st.broadPhaseMs);
            np.push_back(st.narrowPhaseMs);
            aabb.push_back(st.totalMs - st.integrateVelocityMs - st.broadPhaseMs -
                           st.narrowPhaseMs - st.solverMs - st.integratePoseMs);
            total.push_back(st.totalMs);
        }
        auto median = [](std::vector<double> x) {
            if (x.empty())
                return 0.0;
            std::sort(x.begin(), x.end());
            return x[x.size() / 2];
        };
        const StepStats& st = world.lastStepStats();
        if (timing) {
            timing[0] = median(bp);
            timing[1] = median(np);
            timing[2] = median(aabb);
            timing[3] = median(total);
            timing[4] = static_cast<double>(st.candidatePairs);
            timing[5] = static_cast<double>(st.possiblePairs);
        }
        if (npStats) {
            const NarrowPhaseStats& d = st.narrowPhaseDetails;
            npStats[0] = d.sphereSphereCandidates;
            npStats[1] = d.sphereBoxCandidates;
            npStats[2] = d.boxBoxCandidates;
            npStats[3] = d.sphereSphereContacts;
            npStats[4] = d.sphereBoxContacts;
            npStats[5] = d.boxBoxContacts;
        }
        const auto& contacts = world.contacts();
        *count = contacts.size();
        if (out)
            for (size_t k = 0; k < contacts.size() && k < capacity; ++k)
                writeManifold(contacts[k], &out[k]);
        return contacts.size() > capacity ? PHYSGPU_CAPACITY_EXCEEDED : PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int ref_world_collide(const physgpu_colliders* c, const float* bodyPosition,
                      const float* localPose, const float* friction, const float* restitution,
                      int algorithm, uint32_t bpWorkers, uint32_t npWorkers, uint32_t warmup,
                      uint32_t steps, physgpu_manifold* out, uint64_t capacity, uint64_t* count,
                      double* timing, uint64_t* npStats)
{
    return worldCollide(c, bodyPosition, nullptr, localPose, friction, restitution, algorithm,
                        bpWorkers, npWorkers, warmup, steps, out, capacity, count, timing,
                        npStats);
}

// Colliders sharing bodies (colliderBody[i] < n); per-collider arrays carry
// the body's values.  Manifold first/second are body indices (the order in
// which the bodies were added), as ContactManifold::bodyA / bodyB.
int ref_world_collide_bodies(const physgpu_colliders* c, const float* bodyPosition,
                             const uint32_t* colliderBody, const float* localPose,
                             const float* friction, const float* restitution, int algorithm,
                             uint32_t bpWorkers, uint32_t npWorkers, uint32_t warmup,
                             uint32_t steps, physgpu_manifold* out, uint64_t capacity,
                             uint64_t* count, double* timing, uint64_t* npStats)
{
    for (uint32_t i = 0; i < c->count; ++i)
        if (colliderBody[i] >= c->count)
            return fail(std::invalid_argument("body index out of range"));
    return worldCollide(c, bodyPosition, colliderBody, localPose, friction, restitution,
                        algorithm, bpWorkers, npWorkers, warmup, steps, out, capacity, count,
                        timing, npStats);
}

} // extern "C"
