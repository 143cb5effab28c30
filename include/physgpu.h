/*
 * physgpu.h — C ABI of the MI355X (gfx950) collision pipeline for phys-engine.
 *
 * The library replaces the collision section of phys::PhysicsWorld::step
 * (src/world/physics_world.cpp:539-739): AABB refresh, broad phase and narrow
 * phase.  Integration and the sequential-impulse solver stay host code that
 * consumes the manifolds this library returns.
 *
 * Conventions
 *   - No exceptions cross this boundary.  Every call returns a physgpu_status;
 *     physgpu_last_error() gives a thread-local message for the last failure.
 *     Where the reference throws std::invalid_argument the call returns
 *     PHYSGPU_INVALID_ARGUMENT.
 *   - A context is bound to one device and one HIP stream; every call is
 *     stream-ordered.  Calls that return host data synchronise that stream.
 *   - Arrays are plain pointers.  "SoA planes" means plane k of an array with
 *     n entries starts at element k*n.
 *   - Pair indices are 32-bit (the reference uses size_t); ordering is the
 *     reference's canonical lexicographic (first, second) order, first < second.
 */
#ifndef PHYSGPU_H
#define PHYSGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum physgpu_status {
    PHYSGPU_OK = 0,
    PHYSGPU_INVALID_ARGUMENT = 1, /* reference: std::invalid_argument */
    PHYSGPU_CAPACITY_EXCEEDED = 2, /* output or context capacity too small */
    PHYSGPU_DEVICE_ERROR = 3,      /* HIP runtime failure */
    PHYSGPU_UNSUPPORTED = 4,
    /* Every result of the call is written, but some hull pairs got no EPA
     * result (no contact, counted in epa_unresolved); physgpu_epa_unresolved
     * lists them.  Returned by the calls that read the step's counts. */
    PHYSGPU_EPA_UNRESOLVED = 5
} physgpu_status;

/* phys::BroadPhaseAlgorithm, include/phys/collision/broadphase.hpp:8-13.
 * PHYSGPU_DYNAMIC_TREE is served by a linear BVH rebuilt per call. */
typedef enum physgpu_broadphase_algorithm {
    PHYSGPU_SWEEP_AND_PRUNE = 0,
    PHYSGPU_UNIFORM_GRID = 1,
    PHYSGPU_DYNAMIC_TREE = 2
} physgpu_broadphase_algorithm;

/* Alternatives of phys::Collider::shape (collider.hpp:27) plus convex hulls. */
typedef enum physgpu_shape_type {
    PHYSGPU_SPHERE = 0,
    PHYSGPU_BOX = 1,
    PHYSGPU_HULL = 2
} physgpu_shape_type;

/* Where caller arrays live. */
typedef enum physgpu_memory {
    PHYSGPU_HOST = 0,
    PHYSGPU_DEVICE = 1
} physgpu_memory;

/* Same layout as phys::Aabb (include/phys/collision/aabb.hpp:10-13). */
typedef struct physgpu_aabb {
    float min[3];
    float max[3];
} physgpu_aabb;

/* phys::BroadPhasePair (broadphase.hpp:15-19) with 32-bit indices. */
typedef struct physgpu_pair {
    uint32_t first;
    uint32_t second;
} physgpu_pair;

/* A pair of user indices from the persistent proxies (phys::BroadPhasePair
 * with the reference's full-width indices, broadphase.hpp:15-19). */
typedef struct physgpu_proxy_pair {
    uint64_t first;
    uint64_t second;
} physgpu_proxy_pair;

/* phys::BroadPhaseStats (broadphase.hpp:21-42), field for field.  Timings
 * are device time in ms measured with HIP events when timing is enabled on
 * the context, else 0.  After a uniform-grid call the grid_* fields are the
 * reference grid's own counters over the same AABBs (findGridPairs,
 * src/collision/broadphase.cpp:90-172: cell = upper median of the largest
 * sides, entries = cells of the AABBs within 64 cells, comparisons = entry
 * pairs sharing a cell plus the overflow pass), computed on demand when the
 * stats are asked for (a few launches and host waits on the context's
 * stream), not the work of the GPU's own grid.  In a batch of worlds each
 * world is its own grid: the counts are summed and grid_cell_size is the
 * largest world cell. */
typedef struct physgpu_broadphase_stats {
    double record_build_ms;
    double record_sort_ms;
    double sweep_ms;
    double pair_sort_ms;
    uint64_t x_window_comparisons;
    uint64_t aabb_pairs;
    uint64_t grid_comparisons;
    uint64_t grid_entries;
    uint64_t grid_overflow_aabbs;
    double grid_cell_size;
    uint64_t tree_insertions;
    uint64_t tree_removals;
    uint64_t tree_reinsertions;
    uint64_t tree_node_pair_visits;
    uint64_t tree_leaf_checks;
    uint64_t tree_height;
    uint64_t tree_proxy_count;
} physgpu_broadphase_stats;

/* phys::NarrowPhaseStats (step_stats.hpp:7-15) plus the hull classes. */
typedef struct physgpu_narrowphase_stats {
    uint64_t sphere_sphere_candidates;
    uint64_t sphere_box_candidates;
    uint64_t box_box_candidates;
    uint64_t sphere_sphere_contacts;
    uint64_t sphere_box_contacts;
    uint64_t box_box_contacts;
    uint64_t hull_candidates;   /* pairs with at least one hull (GJK/EPA) */
    uint64_t hull_contacts;
    uint64_t gjk_epa_failures;  /* GJK hit but EPA did not converge */
    uint64_t epa_overflows;     /* EPA runs continued past the LDS bounds below */
    uint64_t epa_unresolved;    /* hull pairs left without an EPA result */
} physgpu_narrowphase_stats;

/* The collision fields of phys::StepStats (include/phys/world/
 * step_stats.hpp:33-54) for the last device-resident step (physgpu_collide /
 * physgpu_collide_graph).  Stage times are device time between HIP events
 * around the broad phase (after the AABB refresh) and the narrow phase when
 * timing is enabled on the context (physgpu_set_timing), else 0; the
 * reference's collect and filter sub-stages have no separate device stage
 * (bounds are written by the AABB pass, the same-body filter runs in the
 * narrow phase's classification) and stay 0. */
typedef struct physgpu_step_stats {
    double broad_phase_ms;
    double narrow_phase_ms;
    double broad_phase_collect_ms;
    double broad_phase_filter_ms;
    uint64_t possible_pairs;   /* collider pairs on different bodies, per world
                                  (physics_world.cpp:571-577) */
    uint64_t candidate_pairs;  /* broad-phase pairs left after the same-body
                                  filter (:630-636) */
    uint64_t contact_count;    /* manifolds (:739) */
    physgpu_broadphase_stats broad_phase;
    physgpu_narrowphase_stats narrow_phase;
} physgpu_step_stats;

/* EPA polytope of the GPU narrow phase.  The reference's EPA
 * (src/collision/epa.cpp:64-130) is replayed exactly in LDS while it holds
 * at most PHYSGPU_EPA_MAX_FACES faces and its horizon at most
 * PHYSGPU_EPA_MAX_HORIZON edges.  In FP32 the reference's polytope can fill
 * with copies of the same faces on near-degenerate pairs (hundreds of
 * thousands of faces); a run that outgrows those bounds (counted in
 * epa_overflows) is continued over distinct faces with copy counts, up to
 * PHYSGPU_EPA_MAX_DISTINCT_FACES of them, which gives the reference's result
 * except where the reference's closest-face choice would depend on the order
 * of copies it does not keep: such pairs, and pairs past that bound, get no
 * contact and are counted in epa_unresolved (and listed by
 * physgpu_epa_unresolved). */
#define PHYSGPU_EPA_MAX_FACES 256
#define PHYSGPU_EPA_MAX_HORIZON 128
#define PHYSGPU_EPA_MAX_DISTINCT_FACES 1024
#define PHYSGPU_EPA_MAX_ITERATIONS 64

/* Shape-pair classes.  The reference's shapePair = index(A) + index(B)
 * (physics_world.cpp:670-671) gives 0, 1, 2 for SS, SB/BS, BB; hull classes
 * follow. */
typedef enum physgpu_shape_pair {
    PHYSGPU_PAIR_SPHERE_SPHERE = 0,
    PHYSGPU_PAIR_SPHERE_BOX = 1,
    PHYSGPU_PAIR_BOX_BOX = 2,
    PHYSGPU_PAIR_SPHERE_HULL = 3,
    PHYSGPU_PAIR_BOX_HULL = 4,
    PHYSGPU_PAIR_HULL_HULL = 5
} physgpu_shape_pair;

/* One phys::ContactManifold (manifold.hpp:20-32) without the solver's impulse
 * fields.  Anchors are collider-local for physgpu_generate_contacts (the
 * NarrowPhase::generateContact contract) and body-local after
 * physgpu_set_world (physics_world.cpp:679-686).  160 bytes. */
typedef struct physgpu_manifold {
    uint32_t first;           /* collider index of A (bodyA) */
    uint32_t second;          /* collider index of B (bodyB) */
    float normal[3];          /* from A toward B */
    uint32_t point_count;     /* 1..4 */
    float anchor_a[4][3];
    float anchor_b[4][3];
    float penetration[4];
    float friction;
    float restitution;
    uint32_t pair_index;      /* position in the candidate list */
    uint32_t shape_pair;      /* physgpu_shape_pair */
    uint32_t reserved[2];
} physgpu_manifold;

/* Collider description.  All arrays have `count` entries unless stated.
 *   pose        7 SoA planes: px, py, pz, qx, qy, qz, qw (world transform of
 *               the collider: body pose composed with the local transform)
 *   shape_type  physgpu_shape_type per collider
 *   shape_data  3 SoA planes: sphere (radius, -, -); box half extents;
 *               hull: unused
 *   hull_range  2 SoA planes (first point, point count) into hull_points;
 *               may be NULL when no collider is a hull
 *   hull_points 3 SoA planes of hull_point_count collider-local points */
typedef struct physgpu_colliders {
    uint32_t count;
    const float* pose;
    const uint8_t* shape_type;
    const float* shape_data;
    const uint32_t* hull_range;
    uint32_t hull_point_count;
    const float* hull_points;
} physgpu_colliders;

/* World-level extras for physgpu_set_world, mirroring what
 * PhysicsWorld::step reads besides the collider pose (physics_world.cpp:
 * 570-581, 630-633, 679-690).  Arrays have colliders.count entries; any may be
 * NULL (then: body = collider index, local transform = identity,
 * friction = restitution = 0). */
typedef struct physgpu_world_extras {
    const uint32_t* body;        /* owning body per collider */
    const float* local_pose;     /* 7 SoA planes: collider local transform */
    const float* friction;       /* per collider (its body's value) */
    const float* restitution;    /* per collider (its body's value) */
} physgpu_world_extras;

/* Narrow-phase options.  Zero-initialised means the defaults below. */
typedef struct physgpu_narrowphase_options {
    uint32_t gjk_max_iterations;  /* default 64 */
    uint32_t epa_max_iterations;  /* default 64 */
    float epa_tolerance;          /* default 1e-6 */
    uint32_t reserved;
} physgpu_narrowphase_options;

/* Launch shapes of the tunable kernels.  No counterpart in the reference
 * (CPU code); physgpu_autotune picks them by timing, physgpu_set_tuning sets
 * them.  Zero fields mean the default in brackets. */
typedef struct physgpu_tuning {
    uint32_t class_block[3];   /* workgroup size of the SS, SB/BS, BB kernels: 64, 128, 256 [256] */
    uint32_t class_groups[3];  /* workgroups of those grid-stride kernels: 1..8192 [2048] */
    uint32_t epa_groups;       /* GJK + EPA workgroups of one wavefront; work item e (one
                                  sphere-hull candidate, or four hull candidates) goes to
                                  workgroup e mod epa_groups: 1..1048576 [65536] */
    uint32_t query_block;      /* grid query workgroup size: 64, 128, 256 [256] */
    uint32_t query_groups;     /* [2048] */
    uint32_t sort_items;       /* radix sort items per thread: 4, 8, 16 (tile 256 x items) [8] */
    uint32_t sphere_hull_per_wave; /* sphere-hull candidates per GJK + EPA wavefront: 1 (its EPA
                                      starts at once) or 4 (one GJK per 16-lane row, then their
                                      EPA runs in turn, as for box-hull and hull-hull) [1] */
} physgpu_tuning;

typedef struct physgpu_context physgpu_context;

/* Thread-local text for the last non-OK status on this thread. */
const char* physgpu_last_error(void);

/* Number of HIP devices visible (0 when none); never fails. */
int physgpu_device_count(void);

/* Context on `device` able to hold `max_colliders` colliders,
 * `max_pairs` candidate pairs and `max_hull_points` hull points.  Device
 * buffers are allocated here, once. */
physgpu_status physgpu_create(physgpu_context** out, int device, uint32_t max_colliders,
                              uint64_t max_pairs, uint32_t max_hull_points);
physgpu_status physgpu_destroy(physgpu_context* ctx);

/* Launch subsequent work on `hip_stream` (a hipStream_t; NULL = the null stream). */
physgpu_status physgpu_set_stream(physgpu_context* ctx, void* hip_stream);

/* Enable per-kernel HIP-event timing (stats timings, physgpu_kernel_times). */
physgpu_status physgpu_set_timing(physgpu_context* ctx, int enabled);

/* Copy colliders into the context (memory says where the arrays live). */
physgpu_status physgpu_set_colliders(physgpu_context* ctx, const physgpu_colliders* colliders,
                                     physgpu_memory memory);

/* Replace the poses of the resident colliders (7 SoA planes, one entry per
 * resident collider) and keep their shapes: the per-step input of the AABB
 * refresh (physics_world.cpp:539-551 reads the integrated poses).  A captured
 * step graph stays valid.  Host memory is copied before the call returns;
 * device memory is read in stream order (it must stay unchanged until the
 * context's stream has passed this call). */
physgpu_status physgpu_update_poses(physgpu_context* ctx, const float* pose,
                                    physgpu_memory memory);

/* Aabb::fromCollider (src/collision/aabb.cpp:8-32) for every resident
 * collider; hulls take min/max of their transformed points.  When `out` is not
 * NULL the bounds are copied to it (count entries). */
physgpu_status physgpu_compute_aabbs(physgpu_context* ctx, physgpu_aabb* out);

/* BroadPhase::findCandidatePairs (include/phys/collision/broadphase.hpp:52-54,
 * src/collision/broadphase.cpp:186-291).  `bounds` (host) replaces the
 * resident AABBs when not NULL.  Writes up to `capacity` pairs to `out`
 * (host, may be NULL) and the full count to `*count`; returns
 * PHYSGPU_CAPACITY_EXCEEDED when the pairs do not fit the context.
 * `stats` (may be NULL) is overwritten.  Unlike the reference, the stateless
 * call accepts PHYSGPU_DYNAMIC_TREE (an LBVH is built for the call).
 * `grid_cell_size` <= 0 selects the automatic cell (the largest AABB side,
 * capped at 8x the mean side, widened by 1/256). */
physgpu_status physgpu_find_candidate_pairs(physgpu_context* ctx, const physgpu_aabb* bounds,
                                            uint32_t n, int algorithm, double grid_cell_size,
                                            physgpu_pair* out, uint64_t capacity, uint64_t* count,
                                            physgpu_broadphase_stats* stats);

/* NarrowPhase::generateContact (src/collision/narrowphase.cpp:326-374) for
 * every pair, over the resident colliders.  `pairs` (host) NULL means the
 * candidate pairs of the last broad phase.  Manifolds come back compacted in
 * pair order (collider-local anchors).  `out` may be NULL. */
physgpu_status physgpu_generate_contacts(physgpu_context* ctx, const physgpu_pair* pairs,
                                         uint64_t n_pairs,
                                         const physgpu_narrowphase_options* options,
                                         physgpu_manifold* out, uint64_t capacity,
                                         uint64_t* count, physgpu_narrowphase_stats* stats);

/* The collision section of PhysicsWorld::step, device resident: AABBs,
 * broad phase, narrow phase.  Results stay on the device (read them with the
 * download calls); the two counts are returned.  `grid_cell_size` <= 0
 * selects the automatic cell, here from the shapes (the largest side any
 * collider's AABB can have at any orientation, capped at 8x the mean,
 * widened by 1/256); the pair set does not depend on the cell. */
physgpu_status physgpu_collide(physgpu_context* ctx, int algorithm, double grid_cell_size,
                               const physgpu_narrowphase_options* options,
                               uint64_t* pair_count, uint64_t* manifold_count);

/* World-level form of the collision section (physics_world.cpp:539-739).
 * `body_pose` (7 SoA planes, one entry per resident collider) is the pose of
 * each collider's body; with `extras` the collider's world pose becomes
 * body pose composed with its local transform (physics_world.cpp:660-667),
 * pairs of colliders on the same body are dropped before the narrow phase
 * (:630-633), anchors come back body-local (:679-686) and friction /
 * restitution are combined (:687-690).  Every later physgpu_collide uses
 * these until physgpu_set_colliders is called again. */
physgpu_status physgpu_set_world(physgpu_context* ctx, const float* body_pose,
                                 const physgpu_world_extras* extras, physgpu_memory memory);

physgpu_status physgpu_download_pairs(physgpu_context* ctx, physgpu_pair* out, uint64_t capacity);
physgpu_status physgpu_download_manifolds(physgpu_context* ctx, physgpu_manifold* out,
                                          uint64_t capacity);
/* Counts of the last step: candidate pairs found (may exceed the capacity)
 * and manifolds produced.  Synchronises the context's stream.  Returns
 * PHYSGPU_EPA_UNRESOLVED (counts written) when some hull pairs of the step
 * got no EPA result. */
physgpu_status physgpu_step_counts(physgpu_context* ctx, uint64_t* pair_count,
                                   uint64_t* manifold_count);
/* The hull pairs of the last narrow phase left without an EPA result
 * (epa_unresolved), as (first, second) collider indices in increasing order:
 * up to `capacity` of them to `out` (may be NULL), their number to *count.
 * No counterpart in the reference, whose unbounded list reaches a result
 * (after minutes on some pairs); a caller may run its CPU EPA on these.
 * Synchronises the context's stream. */
physgpu_status physgpu_epa_unresolved(physgpu_context* ctx, physgpu_pair* out, uint64_t capacity,
                                      uint64_t* count);
physgpu_status physgpu_narrowphase_stats_get(physgpu_context* ctx,
                                             physgpu_narrowphase_stats* stats);
/* StepStats collision fields of the last step (see physgpu_step_stats).
 * Synchronises the context's stream. */
physgpu_status physgpu_step_stats_get(physgpu_context* ctx, physgpu_step_stats* stats);

/* ---- Batches of worlds ---------------------------------------------------
 * No counterpart in the reference, which steps one PhysicsWorld at a time;
 * every world keeps the per-world semantics of PhysicsWorld::step
 * (physics_world.cpp:518-782) and yields exactly the pairs and manifolds it
 * would yield alone.  The resident colliders (physgpu_set_colliders) are the
 * concatenation of `world_count` worlds: world w owns colliders
 * [offsets[w], offsets[w + 1]), offsets[0] = 0, offsets[world_count] = the
 * collider count.  Pairs only form inside a world.  Pair and manifold indices
 * stay resident-collider indices; the canonical pair list is the
 * concatenation of the per-world canonical lists (subtract offsets[w] for
 * world-local indices).  world_count = 0 returns to a single world;
 * physgpu_set_colliders also does.  AABBs computed before this call carry
 * no world: run a step (or physgpu_compute_aabbs) before a stateless
 * physgpu_find_candidate_pairs.  While worlds are set the broad phase is
 * PHYSGPU_UNIFORM_GRID (one launch per phase for the whole batch); the other
 * algorithms return PHYSGPU_UNSUPPORTED. */
physgpu_status physgpu_set_worlds(physgpu_context* ctx, uint32_t world_count,
                                  const uint32_t* offsets, physgpu_memory memory);

/* Per-world ranges of the last step: world w's pairs are
 * [pair_offsets[w], pair_offsets[w + 1]) of the canonical pair list and its
 * manifolds [manifold_offsets[w], manifold_offsets[w + 1]) of the manifold
 * list.  Each array has world_count + 1 entries (either may be NULL).
 * Synchronises the context's stream. */
physgpu_status physgpu_world_ranges(physgpu_context* ctx, uint64_t* pair_offsets,
                                    uint64_t* manifold_offsets);

/* physgpu_collide as a HIP graph: the first call captures the device-resident
 * step (AABBs, broad phase, narrow phase, per-world ranges) on the context's
 * stream and later calls replay it.  The graph is captured again after
 * anything that changes the step: colliders, worlds, world extras, stream,
 * tuning, algorithm, cell size or options.  New poses written into the
 * resident buffers between replays (PHYSGPU_DEVICE uploads) are picked up.
 * Needs a stream other than the legacy NULL stream.  Counts as in
 * physgpu_collide (NULL pointers: no synchronisation). */
physgpu_status physgpu_collide_graph(physgpu_context* ctx, int algorithm, double grid_cell_size,
                                     const physgpu_narrowphase_options* options,
                                     uint64_t* pair_count, uint64_t* manifold_count);

/* Launch shapes.  physgpu_autotune times each tunable kernel of the step on
 * the resident colliders (`repeats` steps per candidate, HIP events) for each
 * candidate shape, one kernel at a time, keeps the fastest, and returns the
 * result in `out` (may be NULL).  Results do not depend on the tuning. */
physgpu_status physgpu_set_tuning(physgpu_context* ctx, const physgpu_tuning* tuning);
physgpu_status physgpu_get_tuning(physgpu_context* ctx, physgpu_tuning* tuning);
physgpu_status physgpu_autotune(physgpu_context* ctx, int algorithm, uint32_t repeats,
                                physgpu_tuning* out);

/* ---- Narrow-phase adjoint -------------------------------------------------
 * No counterpart in the reference (which has no gradients); the forward is
 * NarrowPhase::generateContact (narrowphase.cpp:132-324, sat.cpp:38-102).
 * The differentiable outputs of manifold i of the last step (m = the step's
 * manifold count) are its normal, and for k < point_count the depth
 * penetration[k] and the world contact point whose collider-local forms are
 * anchor_a[k] / anchor_b[k].  Supported classes: sphere-sphere, sphere-box /
 * box-sphere and box-box (the SAT axis, clipping vertices or fallback point
 * the forward chose), box-hull and hull-hull (the face EPA converged on, with
 * its vertices' hull points / box corners and the contact point's support
 * features held fixed; GJK and EPA are recomputed).  Sphere-hull manifolds,
 * and hull manifolds whose EPA ran past the LDS polytope bounds (touching
 * contacts of degenerate polytopes), get zero gradients and are counted in
 * *skipped.  Gradients are with respect to the colliders' world poses (the
 * resident poses, or the composed ones after physgpu_set_world); quaternion
 * gradients are tangent to the unit sphere (gradients with respect to the
 * orientation q/|q|).
 *
 * Planes (SoA, m entries each): g_normal 3, g_depth 4 (plane k = point k),
 * g_point 12 (plane 3k + axis); NULL means zero.  Outputs: grad_pose 7 x
 * colliders (px py pz qx qy qz qw, summed over each collider's manifolds;
 * overwritten) and/or grad_pair 14 x m (collider `first`'s 7, then
 * `second`'s); either may be NULL.  `memory` says where all arrays live. */
physgpu_status physgpu_contact_adjoint(physgpu_context* ctx, uint64_t m, const float* g_normal,
                                       const float* g_depth, const float* g_point,
                                       float* grad_pose, float* grad_pair, physgpu_memory memory,
                                       uint64_t* skipped);

/* The differentiable outputs themselves, recomputed by the adjoint's forward
 * (bit-identical normals and depths to the manifolds): normal 3 x m, depth
 * 4 x m, world points 12 x m (zero beyond point_count and for hulls). */
physgpu_status physgpu_contact_geometry(physgpu_context* ctx, uint64_t m, float* normal,
                                        float* depth, float* point, physgpu_memory memory);

/* Persistent proxies: DynamicAabbTree (include/phys/collision/
 * dynamic_aabb_tree.hpp:25-31, src/collision/dynamic_aabb_tree.cpp:108-139,
 * 268-336) with its proxy state on the device.  Each context holds one set of
 * proxies (at most max_colliders); proxy ids are recycled like the
 * reference's.  Bounds and ids are host arrays; every call validates all of
 * its items before changing anything and returns PHYSGPU_INVALID_ARGUMENT
 * (the reference's std::invalid_argument) for non-finite or unordered bounds,
 * unknown or destroyed ids, or a user index already held by a live proxy. */

/* createProxy for `count` proxies; ids to proxy_out. */
physgpu_status physgpu_tree_create_proxies(physgpu_context* ctx, uint32_t count,
                                           const physgpu_aabb* bounds,
                                           const uint64_t* user_index, uint32_t* proxy_out);
/* destroyProxy for `count` proxies. */
physgpu_status physgpu_tree_destroy_proxies(physgpu_context* ctx, uint32_t count,
                                            const uint32_t* proxy);
/* updateProxy for `count` proxies: the tight bounds become `bounds`; the fat
 * bounds follow the reference's rule, and reinserted[i] (may be NULL) is 1
 * where the proxy is reinserted.  An id may appear once per call. */
physgpu_status physgpu_tree_update_proxies(physgpu_context* ctx, uint32_t count,
                                           const uint32_t* proxy, const physgpu_aabb* bounds,
                                           uint8_t* reinserted);
/* Live proxy count (DynamicAabbTree::size). */
physgpu_status physgpu_tree_size(physgpu_context* ctx, uint32_t* live);

typedef enum physgpu_tree_policy {
    /* Build the LBVH over the live proxies' tight bounds for every query. */
    PHYSGPU_TREE_REBUILD = 0,
    /* Keep the last build's leaf order and nodes and recompute the node boxes
     * while no proxy was created or destroyed since that build (and no other
     * broad phase ran on the context); rebuild otherwise. */
    PHYSGPU_TREE_REFIT = 1
} physgpu_tree_policy;

/* DynamicAabbTree::findCandidatePairs: every pair of live proxies whose tight
 * bounds overlap, as (smaller, larger) user index, sorted and unique.  Writes
 * up to `capacity` pairs to `out` (host, may be NULL) and the count to
 * `*count`.  `stats` (may be NULL): aabb_pairs, tree_proxy_count, node visits
 * and leaf checks of the LBVH query, its height, and the creations,
 * destructions and reinsertions since the previous query.  The query uses
 * the context's broad-phase buffers: the resident colliders stay, but the
 * candidate pairs and resident AABBs of earlier calls are replaced, so
 * physgpu_generate_contacts(NULL pairs) and physgpu_find_candidate_pairs(NULL
 * bounds) need a new physgpu_collide / physgpu_compute_aabbs first. */
physgpu_status physgpu_tree_find_pairs(physgpu_context* ctx, int policy, physgpu_proxy_pair* out,
                                       uint64_t capacity, uint64_t* count,
                                       physgpu_broadphase_stats* stats);

/* ---- Allocation-free broad phase -------------------------------------
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  physgpu_find_candidate_pairs over device arrays,
 * with no context: the working memory is one fixed region per device inside
 * the library (room for PHYSGPU_NOALLOC_MAX_BODIES bodies and
 * PHYSGPU_NOALLOC_MAX_PAIRS pairs, about 18 MiB, held on every device the
 * call runs on until the process ends).  Nothing is allocated and nothing
 * synchronises: the work is enqueued on `hip_stream` (a hipStream_t of the
 * current device; NULL = the null stream).
 *   bounds    n device records, finite and ordered (not checked here)
 *   out       device, receives min(count, capacity) pairs in the canonical
 *             order (may be NULL when capacity is 0)
 *   count     device uint64: the full pair count; pairs past
 *             PHYSGPU_NOALLOC_MAX_PAIRS or capacity are not written; ~0 when
 *             the LBVH traversal overflowed (never for valid bounds)
 * Returns PHYSGPU_INVALID_ARGUMENT for n above PHYSGPU_NOALLOC_MAX_BODIES,
 * an unknown algorithm or missing arrays.  All calls on one device share the
 * region and it keeps the last call's contents: calls that overlap in time
 * on one device (other streams, other host threads) race and must be
 * serialised by the caller. */
#define PHYSGPU_NOALLOC_MAX_BODIES 65536
#define PHYSGPU_NOALLOC_MAX_PAIRS 131072
physgpu_status physgpu_find_candidate_pairs_noalloc(const physgpu_aabb* bounds, uint32_t n,
                                                    int algorithm, double grid_cell_size,
                                                    physgpu_pair* out, uint64_t capacity,
                                                    uint64_t* count, void* hip_stream);

/* Average device time (ms) per launch of each kernel since timing was
 * enabled.  Writes up to `capacity` entries; names are static strings. */
physgpu_status physgpu_kernel_times(physgpu_context* ctx, const char** names, double* avg_ms,
                                    uint64_t* launches, uint32_t capacity, uint32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* PHYSGPU_H */
