// C ABI of the collision library (include/physgpu.h): context lifecycle,
// uploads, the stage sequence of PhysicsWorld::step's collision section
// (src/world/physics_world.cpp:539-739), statistics and kernel timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <string>
#include <vector>

#include "internal.hpp"

namespace pg {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }

physgpu_status fail(physgpu_status st, const std::string& msg)
{
    g_error = msg;
    return st;
}

physgpu_status check_hip(hipError_t e, const char* what)
{
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return PHYSGPU_DEVICE_ERROR;
}

const char* const kStageBroadphase = "stage_broadphase";
const char* const kStageNarrowphase = "stage_narrowphase";

void time_begin(physgpu_context* ctx, const char* name, Timer& t, hipStream_t stream)
{
    t.name = name;
    t.start = t.stop = nullptr;
    t.stream = stream ? stream : ctx->stream;
    if (!ctx->timing)
        return;
    auto take = [&]() {
        hipEvent_t e = nullptr;
        if (!ctx->event_pool.empty()) {
            e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
        } else {
            hipEventCreate(&e);
        }
        return e;
    };
    t.start = take();
    t.stop = take();
    hipEventRecord(t.start, t.stream);
}

void time_end(physgpu_context* ctx, Timer& t)
{
    if (!ctx->timing || !t.start)
        return;
    hipEventRecord(t.stop, t.stream);
    ctx->pending.push_back(t);
}

void collect_times(physgpu_context* ctx)
{
    for (Timer& t : ctx->pending) {
        float ms = 0.0f;
        hipEventSynchronize(t.stop);
        hipEventElapsedTime(&ms, t.start, t.stop);
        auto it = std::find_if(ctx->kernel_times.begin(), ctx->kernel_times.end(),
                               [&](const KernelTime& k) { return std::strcmp(k.name, t.name) == 0; });
        if (it == ctx->kernel_times.end()) {
            ctx->kernel_times.push_back({t.name, 0.0, 0});
            it = ctx->kernel_times.end() - 1;
        }
        it->total_ms += ms;
        it->launches += 1;
        if (std::strcmp(t.name, kStageBroadphase) == 0)
            ctx->last_broadphase_ms = ms;
        else if (std::strcmp(t.name, kStageNarrowphase) == 0)
            ctx->last_narrowphase_ms = ms;
        ctx->event_pool.push_back(t.start);
        ctx->event_pool.push_back(t.stop);
    }
    ctx->pending.clear();
}

namespace {

template <typename T>
hipError_t alloc(T*& p, size_t count)
{
    return hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T));
}

void free_all(physgpu_context* c)
{
    void* ptrs[] = {c->pose, c->pose4, c->shape4, c->shape_type, c->shape_data, c->hull_range, c->hull_points,
                    c->aabb, c->sorted_aabb, c->items_a, c->items_b, c->hull_list, c->shape_cell,
                    c->hull_list_count,
                    c->histogram, c->digit_total, c->cell_table, c->overflow_list, c->grid_over_ctr,
                    c->bucket_bounds, c->node_left,
                    c->node_right, c->node_parent, c->node_first, c->node_last, c->node_box, c->level_box,
                    c->pairs_raw, c->pairs_tmp, c->seg_off, c->seg_tiles, c->pair_rank,
                    c->pair_perm, c->sort_scratch, c->sort_scratch_perm,
                    c->seg_lists, c->buckets, c->slots, c->hit, c->tile_sums,
                    c->manifolds, c->body, c->local_pose, c->friction, c->restitution,
                    c->counters, c->world_body_pose, c->world_offsets, c->world_of,
                    c->world_pair_off, c->world_manifold_off, c->adjoint_scratch,
                    c->adjoint_skipped, c->partials, c->tree_tight, c->tree_fat, c->tree_rank,
                    c->tree_rank_user, c->tree_stage_ids, c->tree_stage_bounds,
                    c->tree_stage_flags, c->tree_stage_pairs, c->epa_records,
                    c->epa_unresolved_pairs, c->epa_flags, c->epa_epoch, c->gs_scalars,
                    c->gs_ranges, c->gs_tags, c->gs_owner, c->gs_counts, c->gs_list};
    for (void* p : ptrs)
        if (p)
            hipFree(p);
    if (c->host_counters)
        hipHostFree(c->host_counters);
    if (c->graph_exec)
        hipGraphExecDestroy(c->graph_exec);
    if (c->graph)
        hipGraphDestroy(c->graph);
    for (Timer& t : c->pending) {
        hipEventDestroy(t.start);
        hipEventDestroy(t.stop);
    }
    for (hipEvent_t e : c->event_pool)
        hipEventDestroy(e);
    if (c->sort_event)
        hipEventDestroy(c->sort_event);
    if (c->fork_event)
        hipEventDestroy(c->fork_event);
    if (c->join_event)
        hipEventDestroy(c->join_event);
    if (c->early_join_event)
        hipEventDestroy(c->early_join_event);
    // side_stream and early_stream belong to the device (helper_streams).
    if (c->stream && c->owns_stream)
        hipStreamDestroy(c->stream);
}

// The narrow phase's two helper streams (SS / SB / BB beside GJK + EPA, and
// the EPA hand-over consumer), one pair per device for every context of the
// process: thousands of contexts (a world per context) would otherwise hold
// thousands of streams.  Contexts that share them only queue their helper
// work behind each other; every step still joins its own work by events.
// Created at a device's first context and kept until the process ends.
constexpr int kMaxHelperDevices = 64;
std::mutex g_helper_mutex;
hipStream_t g_helper_side[kMaxHelperDevices] = {};
hipStream_t g_helper_early[kMaxHelperDevices] = {};

hipError_t helper_streams(int device, hipStream_t& side, hipStream_t& early)
{
    if (device < 0 || device >= kMaxHelperDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_helper_mutex);
    hipError_t e = hipSuccess;
    if (!g_helper_side[device])
        e = hipStreamCreateWithFlags(&g_helper_side[device], hipStreamNonBlocking);
    if (e == hipSuccess && !g_helper_early[device])
        e = hipStreamCreateWithFlags(&g_helper_early[device], hipStreamNonBlocking);
    side = g_helper_side[device];
    early = g_helper_early[device];
    return e;
}

physgpu_status read_counters(physgpu_context* ctx)
{
    PG_TRY_HIP(hipMemcpyAsync(ctx->host_counters, ctx->counters, sizeof(Counters),
                              hipMemcpyDeviceToHost, ctx->stream));
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->timing)
        collect_times(ctx);
    return PHYSGPU_OK;
}

void drop_graph(physgpu_context* ctx)
{
    if (ctx->graph_exec)
        hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph)
        hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_algorithm = -1;
}

physgpu_tuning default_tuning()
{
    physgpu_tuning t{};
    for (int c = 0; c < 3; ++c) {
        t.class_block[c] = kBlock;
        t.class_groups[c] = kStrideGrid;
    }
    t.epa_groups = 65536;  // one wavefront each; at 1M bodies (34k entries) 2048 / 16384 / 65536: 638 / 310 / 263 us
    t.query_block = kBlock;
    t.query_groups = 2048;  // at 1M bodies, 256 x 1024 / 2048 / 4096: 71 / 63 / 78 us
    t.sort_items = 8;
    // At 1M bodies the last long sphere-hull EPA runs end the kernel: four to a
    // wavefront, runs queue behind each other (GJK + EPA 292 -> 427 us).
    t.sphere_hull_per_wave = 1;
    return t;
}

bool valid_block(uint32_t b) { return b == 64 || b == 128 || b == 256; }
bool valid_groups(uint32_t g) { return g >= 1 && g <= kMaxPartialBlocks; }
bool valid_epa_groups(uint32_t g) { return g >= 1 && g <= kMaxEpaGroups; }

// Zero fields take the defaults; anything else must be a supported shape.
physgpu_status resolve_tuning(const physgpu_tuning& in, physgpu_tuning& out)
{
    const physgpu_tuning d = default_tuning();
    out = in;
    for (int c = 0; c < 3; ++c) {
        if (!out.class_block[c])
            out.class_block[c] = d.class_block[c];
        if (!out.class_groups[c])
            out.class_groups[c] = d.class_groups[c];
        if (!valid_block(out.class_block[c]) || !valid_groups(out.class_groups[c]))
            return fail(PHYSGPU_INVALID_ARGUMENT, "class kernel shape out of range");
    }
    if (!out.epa_groups)
        out.epa_groups = d.epa_groups;
    if (!out.query_block)
        out.query_block = d.query_block;
    if (!out.query_groups)
        out.query_groups = d.query_groups;
    if (!out.sort_items)
        out.sort_items = d.sort_items;
    if (!out.sphere_hull_per_wave)
        out.sphere_hull_per_wave = d.sphere_hull_per_wave;
    if (!valid_epa_groups(out.epa_groups) || !valid_block(out.query_block) ||
        !valid_groups(out.query_groups))
        return fail(PHYSGPU_INVALID_ARGUMENT, "kernel shape out of range");
    if (out.sort_items != 4 && out.sort_items != 8 && out.sort_items != 16)
        return fail(PHYSGPU_INVALID_ARGUMENT, "sort_items must be 4, 8 or 16");
    if (out.sphere_hull_per_wave != 1 && out.sphere_hull_per_wave != 4)
        return fail(PHYSGPU_INVALID_ARGUMENT, "sphere_hull_per_wave must be 1 or 4");
    return PHYSGPU_OK;
}

// Broad phase on the resident AABB planes (ctx->n bodies).  With defer_sort
// the emitted pairs are left unordered (pairs_raw) for a narrow phase that
// sorts them on its side stream.
// With grid_fused (uniform grid, n >= 2) the AABB refresh of the resident
// colliders is part of the grid's key pass and nothing else needs the width
// statistics.
// With finish_later the pair count is clamped to the capacity by the narrow
// phase that follows (its classification), not by a launch of its own here.
void run_broadphase(physgpu_context* ctx, int algorithm, double cell, bool window_stats,
                    bool defer_sort = false, bool stats_done = false, bool grid_fused = false,
                    bool finish_later = false)
{
    ctx->tree_build_current = false;
    ctx->tree_pairs_loaded = false;
    ctx->np_classified = false;
    if (!stats_done && !grid_fused)
        launch_aabb_reductions(ctx);
    // The broad phase's closing fold (if any) also finishes the pair count.
    ctx->finish_pending = ctx->n >= 2;
    if (ctx->n >= 2) {
        if (algorithm == PHYSGPU_SWEEP_AND_PRUNE) {
            launch_sap_keys(ctx);
            uint64_t* sorted = nullptr;
            radix_sort(ctx, ctx->items_a, ctx->items_b, ctx->n, nullptr, 0, 32, 0, 0, &sorted,
                       ctx->aabb, ctx->sorted_aabb);
            launch_sap_sweep(ctx);
            if (window_stats)
                launch_sap_window_tests(ctx);
        } else if (algorithm == PHYSGPU_UNIFORM_GRID) {
            launch_grid(ctx, cell, grid_fused);
        } else {
            launch_lbvh(ctx);
        }
    }
    if ((ctx->n < 2 || ctx->finish_pending) && !finish_later)
        launch_finish_pairs(ctx);
    ctx->finish_pending = false;
    if (!defer_sort)
        launch_sort_pairs(ctx, ctx->stream);
    ctx->last_algorithm = algorithm;
}

// The grid fields are the reference grid's own counters over the resident
// AABBs (reference_grid_stats), not the work of the GPU's grid.
physgpu_status fill_broad_stats(physgpu_context* ctx, physgpu_broadphase_stats* s)
{
    std::memset(s, 0, sizeof(*s));
    const Counters& c = *ctx->host_counters;
    s->aabb_pairs = c.pair_count;
    if (ctx->last_algorithm == PHYSGPU_SWEEP_AND_PRUNE) {
        s->x_window_comparisons = c.window_tests;
    } else if (ctx->last_algorithm == PHYSGPU_UNIFORM_GRID) {
        return reference_grid_stats(ctx, s);
    } else if (ctx->last_algorithm == PHYSGPU_DYNAMIC_TREE) {
        s->tree_insertions = ctx->n;
        s->tree_node_pair_visits = c.tree_visits;
        s->tree_leaf_checks = c.tree_leaf_checks;
        s->tree_height = c.tree_height;
        s->tree_proxy_count = ctx->n;
    }
    return PHYSGPU_OK;
}

void fill_np_stats(const Counters& c, physgpu_narrowphase_stats* s)
{
    s->sphere_sphere_candidates = c.class_count[0];
    s->sphere_box_candidates = c.class_count[1];
    s->box_box_candidates = c.class_count[2];
    s->sphere_sphere_contacts = c.class_contacts[0];
    s->sphere_box_contacts = c.class_contacts[1];
    s->box_box_contacts = c.class_contacts[2];
    s->hull_candidates = uint64_t(c.class_count[3]) + c.class_count[4] + c.class_count[5];
    s->hull_contacts = uint64_t(c.class_contacts[3]) + c.class_contacts[4] + c.class_contacts[5];
    s->gjk_epa_failures = c.gjk_failures;
    s->epa_overflows = c.epa_overflow;
    s->epa_unresolved = c.epa_unresolved;
}

// The step's results are complete except for hull pairs whose EPA result
// could not be determined (physgpu_epa_unresolved lists them).
physgpu_status unresolved_status(const Counters& c)
{
    if (!c.epa_unresolved)
        return PHYSGPU_OK;
    return fail(PHYSGPU_EPA_UNRESOLVED,
                std::to_string(c.epa_unresolved) +
                    " hull pair(s) without an EPA result (physgpu_epa_unresolved lists them)");
}

physgpu_status apply_options(physgpu_context* ctx, const physgpu_narrowphase_options* o)
{
    physgpu_narrowphase_options d{64, 64, 1e-6f, 0};
    if (o) {
        if (o->gjk_max_iterations)
            d.gjk_max_iterations = o->gjk_max_iterations;
        if (o->epa_max_iterations)
            d.epa_max_iterations = o->epa_max_iterations;
        if (o->epa_tolerance > 0.0f)
            d.epa_tolerance = o->epa_tolerance;
    }
    if (d.epa_max_iterations > PHYSGPU_EPA_MAX_ITERATIONS)
        return fail(PHYSGPU_INVALID_ARGUMENT, "epa_max_iterations above 64 is not supported");
    ctx->np_options = d;
    return PHYSGPU_OK;
}

// Checks shared by the device-resident step entry points.
physgpu_status check_step(physgpu_context* ctx, int algorithm)
{
    if (!ctx->colliders_valid)
        return fail(PHYSGPU_INVALID_ARGUMENT, "no colliders set");
    if (algorithm < PHYSGPU_SWEEP_AND_PRUNE || algorithm > PHYSGPU_DYNAMIC_TREE)
        return fail(PHYSGPU_INVALID_ARGUMENT, "unknown broadphase algorithm");
    if (ctx->n_worlds && algorithm != PHYSGPU_UNIFORM_GRID)
        return fail(PHYSGPU_UNSUPPORTED, "a batch of worlds runs the uniform-grid broad phase");
    return PHYSGPU_OK;
}

// The collision section of PhysicsWorld::step, enqueued on ctx->stream
// without any host synchronisation (so it can be captured into a graph).
void enqueue_step(physgpu_context* ctx, int algorithm, double cell)
{
    // (The AABB launch resets the step counters.)
    if (ctx->has_world)
        launch_compose_world(ctx);
    // The uniform grid computes the AABBs in its key pass (one launch).
    const bool fused = algorithm == PHYSGPU_UNIFORM_GRID && ctx->n >= 2;
    if (!fused)
        launch_compute_aabbs(ctx, true);
    // Stage brackets as StepStats times them (physics_world.cpp:554-636,
    // 640-738): the broad phase after the AABB refresh (for the grid with
    // it), then the narrow phase.
    Timer bp, np;
    time_begin(ctx, kStageBroadphase, bp);
    run_broadphase(ctx, algorithm, cell, false, true, true, fused, true);
    time_end(ctx, bp);
    ctx->np_pairs = ctx->pairs_raw;
    ctx->np_perm = ctx->pair_perm;
    time_begin(ctx, kStageNarrowphase, np);
    launch_narrowphase(ctx, ctx->has_world, true);
    time_end(ctx, np);
    launch_world_ranges(ctx);
}

physgpu_status step_counts_checked(physgpu_context* ctx, uint64_t* pair_count,
                                   uint64_t* manifold_count)
{
    if (!pair_count && !manifold_count)
        return PHYSGPU_OK;
    physgpu_status st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    const Counters& c = *ctx->host_counters;
    if (pair_count)
        *pair_count = c.pair_count;
    if (manifold_count)
        *manifold_count = c.manifold_count;
    if (c.tree_overflow)
        return fail(PHYSGPU_DEVICE_ERROR, "LBVH traversal stack overflow");
    if (c.pair_count > ctx->cap_p)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "candidate pairs exceed the context capacity");
    return unresolved_status(c);
}

physgpu_status copy_pairs_to_host(physgpu_context* ctx, physgpu_pair* out, uint64_t n)
{
    if (!out || n == 0)
        return PHYSGPU_OK;
    std::vector<uint64_t> tmp(n);
    PG_TRY_HIP(hipMemcpyAsync(tmp.data(), ctx->pairs, n * sizeof(uint64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t k = 0; k < n; ++k) {
        out[k].first = uint32_t(tmp[k] >> 32);
        out[k].second = uint32_t(tmp[k]);
    }
    return PHYSGPU_OK;
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" {

const char* physgpu_last_error(void) { return g_error.c_str(); }

int physgpu_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

physgpu_status physgpu_create(physgpu_context** out, int device, uint32_t max_colliders,
                              uint64_t max_pairs, uint32_t max_hull_points)
{
    if (!out)
        return fail(PHYSGPU_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (max_colliders == 0 || max_colliders > (1u << 30))
        return fail(PHYSGPU_INVALID_ARGUMENT, "max_colliders must be in [1, 2^30]");
    if (max_pairs == 0 || max_pairs > 0xFFFFFFFFull)
        return fail(PHYSGPU_INVALID_ARGUMENT, "max_pairs must be in [1, 2^32)");
    int count = physgpu_device_count();
    if (device < 0 || device >= count)
        return fail(PHYSGPU_INVALID_ARGUMENT, "no such HIP device");
    PG_TRY_HIP(hipSetDevice(device));
    physgpu_context* c = new (std::nothrow) physgpu_context();
    if (!c)
        return fail(PHYSGPU_DEVICE_ERROR, "out of host memory");
    c->device = device;
    c->tuning = default_tuning();
    c->cap_n = max_colliders;
    c->cap_p = max_pairs;
    c->cap_h = max_hull_points;
    c->table_bits = grid_table_bits(max_colliders);
    const size_t N = max_colliders, P = max_pairs;
    const size_t tiles = std::max<size_t>(div_up(N, kSortMinTile), div_up(P, kSortMinTile));
    c->hist_capacity = uint32_t(256 * tiles);
    hipError_t e = hipSuccess;
    auto A = [&](hipError_t r) {
        if (e == hipSuccess)
            e = r;
    };
    A(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->owns_stream = true;
    A(helper_streams(device, c->side_stream, c->early_stream));
    A(hipEventCreateWithFlags(&c->sort_event, hipEventDisableTiming));
    A(hipEventCreateWithFlags(&c->fork_event, hipEventDisableTiming));
    A(hipEventCreateWithFlags(&c->join_event, hipEventDisableTiming));
    A(hipEventCreateWithFlags(&c->early_join_event, hipEventDisableTiming));
    A(alloc(c->pose, 7 * N));
    A(alloc(c->pose4, 2 * N));
    A(alloc(c->shape4, 2 * N));
    A(alloc(c->world_body_pose, 7 * N));
    A(alloc(c->shape_type, N));
    A(alloc(c->shape_data, 3 * N));
    A(alloc(c->hull_range, 2 * N));
    A(alloc(c->hull_points, 3 * size_t(std::max<uint32_t>(max_hull_points, 1))));
    A(alloc(c->hull_list, N));
    A(alloc(c->shape_cell, 4));
    A(alloc(c->hull_list_count, 1));
    A(alloc(c->aabb, 2 * N));
    A(alloc(c->sorted_aabb, 2 * N));
    A(alloc(c->items_a, N));
    A(alloc(c->items_b, N));
    A(alloc(c->histogram, c->hist_capacity));
    A(alloc(c->digit_total, 256));
    A(alloc(c->cell_table, 2 * (size_t(1) << c->table_bits)));
    A(alloc(c->overflow_list, N));
    A(alloc(c->grid_over_ctr, 1));
    A(alloc(c->bucket_bounds, 2 * kMaxGridBuckets));
    A(alloc(c->node_left, N));
    A(alloc(c->node_right, N));
    A(alloc(c->node_parent, 2 * N));
    A(alloc(c->node_last, N));
    A(alloc(c->node_box, 2 * N));
    A(alloc(c->node_first, N));
    A(alloc(c->level_box, 2 * lbvh_level_boxes(N)));
    A(alloc(c->pairs_raw, P));
    A(alloc(c->pairs_tmp, P));
    A(alloc(c->seg_off, N + 1));
    A(alloc(c->seg_tiles, div_up(N + 1, 2048)));
    A(alloc(c->pair_rank, P));
    A(alloc(c->pair_perm, P));
    A(alloc(c->sort_scratch, P));
    A(alloc(c->sort_scratch_perm, P));
    A(alloc(c->seg_lists, 2 * N));
    A(alloc(c->buckets, kClasses * P));
    A(alloc(c->slots, P));
    A(alloc(c->hit, P));
    A(alloc(c->tile_sums, div_up(P, kScanTile) + 1));
    A(alloc(c->manifolds, P));
    A(alloc(c->body, N));
    A(alloc(c->local_pose, 7 * N));
    A(alloc(c->friction, N));
    A(alloc(c->restitution, N));
    A(alloc(c->world_of, N));
    A(alloc(c->adjoint_skipped, 1));
    A(alloc(c->partials, size_t(kPartialSlots) * kMaxPartialBlocks));
    // EPA hand-over records: one per thousand candidate pairs (a few per
    // thousand hull pairs outgrow the LDS polytope in dense mixed scenes).
    c->epa_record_cap = uint32_t(std::min<uint64_t>(std::max<uint64_t>(P / 1024, 64), 8192));
    A(alloc(c->epa_records, size_t(c->epa_record_cap) * kEpaRecordWords));
    A(alloc(c->epa_unresolved_pairs, c->epa_record_cap));
    A(alloc(c->epa_flags, 2 * size_t(c->epa_record_cap) + 2));
    A(alloc(c->epa_epoch, 1));
    // Hand-over flags compare with the epoch of each narrow phase (from 1).
    if (e == hipSuccess)
        A(hipMemsetAsync(c->epa_flags, 0, (2 * size_t(c->epa_record_cap) + 2) * 4, c->stream));
    if (e == hipSuccess)
        A(hipMemsetAsync(c->epa_epoch, 0, 4, c->stream));
    if (e == hipSuccess)
        A(hipMemsetAsync(c->grid_over_ctr, 0, 4, c->stream));

    A(alloc(c->counters, 1));
    A(hipHostMalloc(reinterpret_cast<void**>(&c->host_counters), sizeof(Counters)));
    if (e != hipSuccess) {
        free_all(c);
        delete c;
        return check_hip(e, "physgpu_create allocation");
    }
    std::memset(c->host_counters, 0, sizeof(Counters));
    c->pairs = c->pairs_raw;
    launch_step_begin(c);
    PG_TRY_HIP(hipStreamSynchronize(c->stream));
    c->np_options = {64, 64, 1e-6f, 0};
    *out = c;
    return PHYSGPU_OK;
}

physgpu_status physgpu_destroy(physgpu_context* ctx)
{
    if (!ctx)
        return PHYSGPU_OK;
    hipSetDevice(ctx->device);
    if (ctx->stream)
        hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return PHYSGPU_OK;
}

physgpu_status physgpu_set_stream(physgpu_context* ctx, void* hip_stream)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    if (ctx->owns_stream)
        hipStreamDestroy(ctx->stream);
    ctx->stream = static_cast<hipStream_t>(hip_stream);
    ctx->owns_stream = false;
    return PHYSGPU_OK;
}

physgpu_status physgpu_set_timing(physgpu_context* ctx, int enabled)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    collect_times(ctx);
    ctx->timing = enabled != 0;
    ctx->kernel_times.clear();
    return PHYSGPU_OK;
}

physgpu_status physgpu_set_colliders(physgpu_context* ctx, const physgpu_colliders* col,
                                     physgpu_memory memory)
{
    if (!ctx || !col)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or colliders is NULL");
    const uint32_t n = col->count;
    if (n > ctx->cap_n)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "more colliders than the context holds");
    if (n && (!col->pose || !col->shape_type || !col->shape_data))
        return fail(PHYSGPU_INVALID_ARGUMENT, "pose, shape_type and shape_data are required");
    if (col->hull_point_count > ctx->cap_h)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "more hull points than the context holds");
    bool hulls = col->hull_range != nullptr;
    if (memory == PHYSGPU_HOST) {
        hulls = false;
        for (uint32_t i = 0; i < n; ++i) {
            const uint8_t t = col->shape_type[i];
            if (t > PHYSGPU_HULL)
                return fail(PHYSGPU_INVALID_ARGUMENT, "unknown shape type");
            if (t == PHYSGPU_HULL) {
                hulls = true;
                if (!col->hull_range || !col->hull_points)
                    return fail(PHYSGPU_INVALID_ARGUMENT, "hull collider without hull data");
                const uint64_t first = col->hull_range[i], cnt = col->hull_range[n + i];
                if (cnt == 0 || first + cnt > col->hull_point_count)
                    return fail(PHYSGPU_INVALID_ARGUMENT, "hull_range outside hull_points");
            }
        }
    } else if (memory != PHYSGPU_DEVICE) {
        return fail(PHYSGPU_INVALID_ARGUMENT, "unknown memory kind");
    }
    PG_TRY_HIP(hipSetDevice(ctx->device));
    const hipMemcpyKind kind =
        memory == PHYSGPU_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    if (n) {
        PG_TRY_HIP(hipMemcpyAsync(ctx->pose, col->pose, 7ull * n * 4, kind, ctx->stream));
        PG_TRY_HIP(hipMemcpyAsync(ctx->shape_type, col->shape_type, n, kind, ctx->stream));
        PG_TRY_HIP(hipMemcpyAsync(ctx->shape_data, col->shape_data, 3ull * n * 4, kind, ctx->stream));
        if (col->hull_range)
            PG_TRY_HIP(hipMemcpyAsync(ctx->hull_range, col->hull_range, 2ull * n * 4, kind,
                                      ctx->stream));
        if (col->hull_points && col->hull_point_count)
            PG_TRY_HIP(hipMemcpyAsync(ctx->hull_points, col->hull_points,
                                      3ull * col->hull_point_count * 4, kind, ctx->stream));
    }
    ctx->n = n;
    ctx->n_hull_points = col->hull_point_count;
    launch_pose_aos(ctx);
    launch_shape_aos(ctx);
    if (hulls)
        launch_hull_list(ctx);
    launch_shape_cell(ctx);
    PG_TRY_HIP(hipGetLastError());
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    ctx->has_hulls = hulls;
    ctx->colliders_valid = true;
    ctx->has_world = false;
    ctx->n_worlds = 0;
    drop_graph(ctx);
    return PHYSGPU_OK;
}

physgpu_status physgpu_update_poses(physgpu_context* ctx, const float* pose,
                                    physgpu_memory memory)
{
    if (!ctx || !pose)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or pose is NULL");
    if (!ctx->colliders_valid)
        return fail(PHYSGPU_INVALID_ARGUMENT, "no colliders set");
    if (memory != PHYSGPU_HOST && memory != PHYSGPU_DEVICE)
        return fail(PHYSGPU_INVALID_ARGUMENT, "unknown memory kind");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    if (memory == PHYSGPU_HOST) {
        if (ctx->n)
            PG_TRY_HIP(hipMemcpyAsync(ctx->pose, pose, 7ull * ctx->n * 4, hipMemcpyHostToDevice,
                                      ctx->stream));
        launch_pose_aos(ctx);
        PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        // Device planes are read in place by the layout kernel (stream order).
        launch_pose_aos(ctx, pose);
    }
    PG_TRY_HIP(hipGetLastError());
    return PHYSGPU_OK;
}

physgpu_status physgpu_compute_aabbs(physgpu_context* ctx, physgpu_aabb* out)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    if (!ctx->colliders_valid)
        return fail(PHYSGPU_INVALID_ARGUMENT, "no colliders set");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    launch_compute_aabbs(ctx);
    if (out && ctx->n) {
        physgpu_aabb* staging = reinterpret_cast<physgpu_aabb*>(ctx->sorted_aabb);
        launch_convert_aabbs_out(ctx, staging);
        PG_TRY_HIP(hipMemcpyAsync(out, staging, sizeof(physgpu_aabb) * ctx->n,
                                  hipMemcpyDeviceToHost, ctx->stream));
    }
    PG_TRY_HIP(hipGetLastError());
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return PHYSGPU_OK;
}

physgpu_status physgpu_find_candidate_pairs(physgpu_context* ctx, const physgpu_aabb* bounds,
                                            uint32_t n, int algorithm, double grid_cell_size,
                                            physgpu_pair* out, uint64_t capacity, uint64_t* count,
                                            physgpu_broadphase_stats* stats)
{
    if (!ctx || !count)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or count is NULL");
    if (algorithm < PHYSGPU_SWEEP_AND_PRUNE || algorithm > PHYSGPU_DYNAMIC_TREE)
        return fail(PHYSGPU_INVALID_ARGUMENT, "unknown broadphase algorithm");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    if (bounds) {
        if (n > ctx->cap_n)
            return fail(PHYSGPU_CAPACITY_EXCEEDED, "more bounds than the context holds");
        for (uint32_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a)
                if (!(bounds[i].min[a] <= bounds[i].max[a]) ||
                    !(bounds[i].min[a] >= -3.402823466e+38f) ||
                    !(bounds[i].max[a] <= 3.402823466e+38f))
                    return fail(PHYSGPU_INVALID_ARGUMENT, "bounds must be finite and ordered");
        ctx->n = n;
        ctx->colliders_valid = false;
        ctx->has_world = false;
        ctx->n_worlds = 0;
        drop_graph(ctx);
        if (n) {
            physgpu_aabb* staging = reinterpret_cast<physgpu_aabb*>(ctx->sorted_aabb);
            PG_TRY_HIP(hipMemcpyAsync(staging, bounds, sizeof(physgpu_aabb) * n,
                                      hipMemcpyHostToDevice, ctx->stream));
            launch_convert_aabbs_in(ctx, staging);
        }
        ctx->resident_aabbs_valid = true;
    } else if (n != ctx->n) {
        return fail(PHYSGPU_INVALID_ARGUMENT, "n must equal the resident collider count");
    } else if (!ctx->resident_aabbs_valid) {
        return fail(PHYSGPU_INVALID_ARGUMENT,
                    "a proxy-tree query replaced the resident AABBs: compute them again first");
    } else if (ctx->n_worlds && algorithm != PHYSGPU_UNIFORM_GRID) {
        return fail(PHYSGPU_UNSUPPORTED, "a batch of worlds runs the uniform-grid broad phase");
    }
    launch_step_begin(ctx);
    run_broadphase(ctx, algorithm, grid_cell_size, stats != nullptr);
    PG_TRY_HIP(hipGetLastError());
    physgpu_status st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    const Counters& c = *ctx->host_counters;
    *count = c.pair_count;
    if (stats) {
        st = fill_broad_stats(ctx, stats);
        if (st != PHYSGPU_OK)
            return st;
    }
    if (c.tree_overflow)
        return fail(PHYSGPU_DEVICE_ERROR, "LBVH traversal stack overflow");
    if (c.pair_count > ctx->cap_p)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "candidate pairs exceed the context capacity");
    st = copy_pairs_to_host(ctx, out, std::min<uint64_t>(capacity, c.pair_count));
    if (st != PHYSGPU_OK)
        return st;
    if (out && capacity < c.pair_count)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "output capacity below the pair count");
    return PHYSGPU_OK;
}

physgpu_status physgpu_generate_contacts(physgpu_context* ctx, const physgpu_pair* pairs,
                                         uint64_t n_pairs,
                                         const physgpu_narrowphase_options* options,
                                         physgpu_manifold* out, uint64_t capacity,
                                         uint64_t* count, physgpu_narrowphase_stats* stats)
{
    if (!ctx || !count)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or count is NULL");
    if (!ctx->colliders_valid)
        return fail(PHYSGPU_INVALID_ARGUMENT, "no colliders set");
    physgpu_status st = apply_options(ctx, options);
    if (st != PHYSGPU_OK)
        return st;
    PG_TRY_HIP(hipSetDevice(ctx->device));
    if (pairs) {
        if (n_pairs > ctx->cap_p)
            return fail(PHYSGPU_CAPACITY_EXCEEDED, "more pairs than the context holds");
        for (uint64_t k = 0; k < n_pairs; ++k)
            if (pairs[k].first >= ctx->n || pairs[k].second >= ctx->n)
                return fail(PHYSGPU_INVALID_ARGUMENT, "pair index out of range");
        launch_step_begin(ctx);
        if (n_pairs)
            PG_TRY_HIP(hipMemcpyAsync(ctx->pairs_raw, pairs, n_pairs * sizeof(physgpu_pair),
                                      hipMemcpyHostToDevice, ctx->stream));
        launch_pairs_in(ctx, nullptr, n_pairs);
    } else {
        if (ctx->tree_pairs_loaded)
            return fail(PHYSGPU_INVALID_ARGUMENT,
                        "the last broad phase was a proxy-tree query: pass pairs or collide first");
        launch_np_begin(ctx);
    }
    // Given pairs come back in their own order, the last broad phase's in
    // canonical order.
    ctx->np_pairs = ctx->pairs;
    ctx->np_perm = nullptr;
    launch_narrowphase(ctx, false, false);
    PG_TRY_HIP(hipGetLastError());
    st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    const Counters& c = *ctx->host_counters;
    *count = c.manifold_count;
    if (stats)
        fill_np_stats(c, stats);
    if (out && c.manifold_count) {
        const uint64_t m = std::min<uint64_t>(capacity, c.manifold_count);
        PG_TRY_HIP(hipMemcpyAsync(out, ctx->manifolds, m * sizeof(physgpu_manifold),
                                  hipMemcpyDeviceToHost, ctx->stream));
        PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
        if (m < c.manifold_count)
            return fail(PHYSGPU_CAPACITY_EXCEEDED, "output capacity below the manifold count");
    }
    return unresolved_status(c);
}

physgpu_status physgpu_collide(physgpu_context* ctx, int algorithm, double grid_cell_size,
                               const physgpu_narrowphase_options* options, uint64_t* pair_count,
                               uint64_t* manifold_count)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    physgpu_status st = check_step(ctx, algorithm);
    if (st != PHYSGPU_OK)
        return st;
    st = apply_options(ctx, options);
    if (st != PHYSGPU_OK)
        return st;
    PG_TRY_HIP(hipSetDevice(ctx->device));
    enqueue_step(ctx, algorithm, grid_cell_size);
    PG_TRY_HIP(hipGetLastError());
    return step_counts_checked(ctx, pair_count, manifold_count);
}

physgpu_status physgpu_collide_graph(physgpu_context* ctx, int algorithm, double grid_cell_size,
                                     const physgpu_narrowphase_options* options,
                                     uint64_t* pair_count, uint64_t* manifold_count)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    physgpu_status st = check_step(ctx, algorithm);
    if (st != PHYSGPU_OK)
        return st;
    st = apply_options(ctx, options);
    if (st != PHYSGPU_OK)
        return st;
    if (!ctx->stream)
        return fail(PHYSGPU_UNSUPPORTED, "graph capture needs a stream other than the NULL stream");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    const bool same = ctx->graph_exec && ctx->graph_algorithm == algorithm &&
                      ctx->graph_cell == grid_cell_size &&
                      std::memcmp(&ctx->graph_options, &ctx->np_options,
                                  sizeof(physgpu_narrowphase_options)) == 0;
    if (!same) {
        drop_graph(ctx);
        // Per-kernel events are not recorded inside the graph.
        const bool timing = ctx->timing;
        ctx->timing = false;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            enqueue_step(ctx, algorithm, grid_cell_size);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(ctx->stream, &g);
            ctx->graph = g;
        }
        ctx->timing = timing;
        if (e == hipSuccess)
            e = hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph(ctx);
            return check_hip(e, "graph capture of the collision step");
        }
        ctx->graph_algorithm = algorithm;
        ctx->graph_cell = grid_cell_size;
        ctx->graph_options = ctx->np_options;
    }
    PG_TRY_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
    ctx->last_algorithm = algorithm;
    return step_counts_checked(ctx, pair_count, manifold_count);
}

physgpu_status physgpu_set_worlds(physgpu_context* ctx, uint32_t world_count,
                                  const uint32_t* offsets, physgpu_memory memory)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    if (!ctx->colliders_valid)
        return fail(PHYSGPU_INVALID_ARGUMENT, "set colliders first");
    drop_graph(ctx);
    if (world_count == 0) {
        ctx->n_worlds = 0;
        ctx->host_world_offsets.clear();
        return PHYSGPU_OK;
    }
    if (!offsets)
        return fail(PHYSGPU_INVALID_ARGUMENT, "offsets is NULL");
    if (memory != PHYSGPU_HOST && memory != PHYSGPU_DEVICE)
        return fail(PHYSGPU_INVALID_ARGUMENT, "unknown memory kind");
    if (world_count > ctx->cap_n)
        return fail(PHYSGPU_INVALID_ARGUMENT, "more worlds than colliders the context holds");
    std::vector<uint32_t> host;
    if (memory == PHYSGPU_DEVICE) {
        host.resize(size_t(world_count) + 1);
        PG_TRY_HIP(hipMemcpy(host.data(), offsets, host.size() * 4, hipMemcpyDeviceToHost));
    } else {
        host.assign(offsets, offsets + size_t(world_count) + 1);
    }
    if (host[0] != 0 || host[world_count] != ctx->n)
        return fail(PHYSGPU_INVALID_ARGUMENT,
                    "offsets must start at 0 and end at the resident collider count");
    for (uint32_t w = 0; w < world_count; ++w)
        if (host[w + 1] < host[w])
            return fail(PHYSGPU_INVALID_ARGUMENT, "offsets must be non-decreasing");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    if (world_count + 1 > ctx->world_capacity) {
        PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
        for (void* p : {(void*)ctx->world_offsets, (void*)ctx->world_pair_off,
                        (void*)ctx->world_manifold_off})
            if (p)
                hipFree(p);
        ctx->world_offsets = nullptr;
        ctx->world_pair_off = ctx->world_manifold_off = nullptr;
        ctx->world_capacity = 0;
        const size_t cap = size_t(world_count) + 1;
        PG_TRY_HIP(alloc(ctx->world_offsets, cap));
        PG_TRY_HIP(alloc(ctx->world_pair_off, cap));
        PG_TRY_HIP(alloc(ctx->world_manifold_off, cap));
        ctx->world_capacity = uint32_t(cap);
    }
    PG_TRY_HIP(hipMemcpyAsync(ctx->world_offsets, host.data(), host.size() * 4,
                              hipMemcpyHostToDevice, ctx->stream));
    ctx->n_worlds = world_count;
    ctx->host_world_offsets = host;
    launch_world_of(ctx);
    PG_TRY_HIP(hipGetLastError());
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return PHYSGPU_OK;
}

physgpu_status physgpu_world_ranges(physgpu_context* ctx, uint64_t* pair_offsets,
                                    uint64_t* manifold_offsets)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    if (!ctx->n_worlds)
        return fail(PHYSGPU_INVALID_ARGUMENT, "no worlds set");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    const size_t bytes = (size_t(ctx->n_worlds) + 1) * sizeof(uint64_t);
    if (pair_offsets)
        PG_TRY_HIP(hipMemcpyAsync(pair_offsets, ctx->world_pair_off, bytes, hipMemcpyDeviceToHost,
                                  ctx->stream));
    if (manifold_offsets)
        PG_TRY_HIP(hipMemcpyAsync(manifold_offsets, ctx->world_manifold_off, bytes,
                                  hipMemcpyDeviceToHost, ctx->stream));
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    return PHYSGPU_OK;
}

physgpu_status physgpu_set_tuning(physgpu_context* ctx, const physgpu_tuning* tuning)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    physgpu_tuning t = default_tuning();
    if (tuning) {
        physgpu_status st = resolve_tuning(*tuning, t);
        if (st != PHYSGPU_OK)
            return st;
    }
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    ctx->tuning = t;
    return PHYSGPU_OK;
}

physgpu_status physgpu_get_tuning(physgpu_context* ctx, physgpu_tuning* tuning)
{
    if (!ctx || !tuning)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or tuning is NULL");
    *tuning = ctx->tuning;
    return PHYSGPU_OK;
}

physgpu_status physgpu_autotune(physgpu_context* ctx, int algorithm, uint32_t repeats,
                                physgpu_tuning* out)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    physgpu_status st = check_step(ctx, algorithm);
    if (st != PHYSGPU_OK)
        return st;
    if (repeats == 0)
        repeats = 3;
    PG_TRY_HIP(hipSetDevice(ctx->device));
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    drop_graph(ctx);
    const bool timing = ctx->timing;
    collect_times(ctx);
    const std::vector<KernelTime> saved = ctx->kernel_times;
    ctx->timing = true;
    // Average device time of the named kernels over `repeats` steps.
    auto measure = [&](std::initializer_list<const char*> names, double& ms) -> physgpu_status {
        enqueue_step(ctx, algorithm, 0.0);  // warm-up
        PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
        collect_times(ctx);
        ctx->kernel_times.clear();
        for (uint32_t r = 0; r < repeats; ++r)
            enqueue_step(ctx, algorithm, 0.0);
        PG_TRY_HIP(hipGetLastError());
        PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
        collect_times(ctx);
        ms = 0.0;
        for (const KernelTime& k : ctx->kernel_times)
            for (const char* nm : names)
                if (std::strcmp(k.name, nm) == 0)
                    ms += k.total_ms / double(repeats);
        return PHYSGPU_OK;
    };
    static const uint32_t blocks[] = {64, 128, 256};
    static const uint32_t groups[] = {256, 512, 1024, 2048, 4096, 8192};
    // One kernel at a time: try each (block, groups) shape, keep the fastest.
    auto tune_shape = [&](uint32_t* block, uint32_t* group, const char* name) -> physgpu_status {
        uint32_t best_b = *block, best_g = *group;
        double best = 1e30;
        for (uint32_t b : blocks)
            for (uint32_t g : groups) {
                if (block)
                    *block = b;
                *group = g;
                double ms = 0.0;
                physgpu_status s = measure({name}, ms);
                if (s != PHYSGPU_OK)
                    return s;
                if (ms < best) {
                    best = ms;
                    best_b = b;
                    best_g = g;
                }
            }
        *block = best_b;
        *group = best_g;
        return PHYSGPU_OK;
    };
    static const char* class_names[3] = {"np_sphere_sphere", "np_sphere_box", "np_box_box"};
    for (int c = 0; c < 3 && st == PHYSGPU_OK; ++c)
        st = tune_shape(&ctx->tuning.class_block[c], &ctx->tuning.class_groups[c], class_names[c]);
    if (st == PHYSGPU_OK && ctx->has_hulls) {
        static const uint32_t epa_groups[] = {4096, 8192, 16384, 32768, 65536};
        uint32_t best_g = ctx->tuning.epa_groups;
        double best = 1e30;
        for (uint32_t g : epa_groups) {
            ctx->tuning.epa_groups = g;
            double ms = 0.0;
            st = measure({"np_gjk_epa"}, ms);
            if (st != PHYSGPU_OK)
                break;
            if (ms < best) {
                best = ms;
                best_g = g;
            }
        }
        ctx->tuning.epa_groups = best_g;
        uint32_t best_w = ctx->tuning.sphere_hull_per_wave;
        best = 1e30;
        for (uint32_t w : {1u, 4u}) {
            if (st != PHYSGPU_OK)
                break;
            ctx->tuning.sphere_hull_per_wave = w;
            double ms = 0.0;
            st = measure({"np_gjk_epa"}, ms);
            if (st == PHYSGPU_OK && ms < best) {
                best = ms;
                best_w = w;
            }
        }
        ctx->tuning.sphere_hull_per_wave = best_w;
    }
    if (st == PHYSGPU_OK && algorithm == PHYSGPU_UNIFORM_GRID)
        st = tune_shape(&ctx->tuning.query_block, &ctx->tuning.query_groups, "grid_query");
    if (st == PHYSGPU_OK) {
        static const uint32_t items[] = {4, 8, 16};
        uint32_t best_i = ctx->tuning.sort_items;
        double best = 1e30;
        for (uint32_t it : items) {
            ctx->tuning.sort_items = it;
            double ms = 0.0;
            st = measure({"radix_upsweep", "radix_scan", "radix_downsweep"}, ms);
            if (st != PHYSGPU_OK)
                break;
            if (ms < best) {
                best = ms;
                best_i = it;
            }
        }
        ctx->tuning.sort_items = best_i;
    }
    ctx->timing = timing;
    ctx->kernel_times = saved;
    if (st != PHYSGPU_OK)
        return st;
    if (out)
        *out = ctx->tuning;
    return PHYSGPU_OK;
}

physgpu_status physgpu_set_world(physgpu_context* ctx, const float* body_pose,
                                 const physgpu_world_extras* extras, physgpu_memory memory)
{
    if (!ctx || !body_pose)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or body_pose is NULL");
    if (!ctx->colliders_valid)
        return fail(PHYSGPU_INVALID_ARGUMENT, "set colliders first");
    const uint32_t n = ctx->n;
    if (memory == PHYSGPU_HOST && extras && extras->body)
        for (uint32_t i = 0; i < n; ++i)
            if (extras->body[i] >= n)
                return fail(PHYSGPU_INVALID_ARGUMENT, "body index out of range");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    const hipMemcpyKind kind =
        memory == PHYSGPU_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    if (n) {
        PG_TRY_HIP(hipMemcpyAsync(ctx->world_body_pose, body_pose, 7ull * n * 4, kind, ctx->stream));
        const bool has_local = extras && extras->local_pose;
        if (has_local)
            PG_TRY_HIP(hipMemcpyAsync(ctx->local_pose, extras->local_pose, 7ull * n * 4, kind,
                                      ctx->stream));
        if (extras && extras->body)
            PG_TRY_HIP(hipMemcpyAsync(ctx->body, extras->body, 4ull * n, kind, ctx->stream));
        if (extras && extras->friction)
            PG_TRY_HIP(hipMemcpyAsync(ctx->friction, extras->friction, 4ull * n, kind, ctx->stream));
        if (extras && extras->restitution)
            PG_TRY_HIP(hipMemcpyAsync(ctx->restitution, extras->restitution, 4ull * n, kind,
                                      ctx->stream));
        ctx->world_has_local = has_local;
        ctx->world_has_body = extras && extras->body;
        ctx->host_body.clear();
        if (ctx->world_has_body) {
            ctx->host_body.resize(n);
            PG_TRY_HIP(hipMemcpyAsync(ctx->host_body.data(), ctx->body, 4ull * n,
                                      hipMemcpyDeviceToHost, ctx->stream));
        }
        ctx->world_has_friction = extras && extras->friction;
        ctx->world_has_restitution = extras && extras->restitution;
    }
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    ctx->has_world = true;
    drop_graph(ctx);
    return PHYSGPU_OK;
}

physgpu_status physgpu_download_pairs(physgpu_context* ctx, physgpu_pair* out, uint64_t capacity)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    physgpu_status st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    const uint64_t held = ctx->host_counters->sort_count;
    st = copy_pairs_to_host(ctx, out, std::min<uint64_t>(capacity, held));
    if (st != PHYSGPU_OK)
        return st;
    if (capacity < ctx->host_counters->pair_count)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "capacity below the pair count");
    return PHYSGPU_OK;
}

physgpu_status physgpu_download_manifolds(physgpu_context* ctx, physgpu_manifold* out,
                                          uint64_t capacity)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    physgpu_status st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    const uint64_t m = std::min<uint64_t>(capacity, ctx->host_counters->manifold_count);
    if (out && m) {
        PG_TRY_HIP(hipMemcpyAsync(out, ctx->manifolds, m * sizeof(physgpu_manifold),
                                  hipMemcpyDeviceToHost, ctx->stream));
        PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (capacity < ctx->host_counters->manifold_count)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "capacity below the manifold count");
    return PHYSGPU_OK;
}

physgpu_status physgpu_step_counts(physgpu_context* ctx, uint64_t* pair_count,
                                   uint64_t* manifold_count)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    physgpu_status st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    if (pair_count)
        *pair_count = ctx->host_counters->pair_count;
    if (manifold_count)
        *manifold_count = ctx->host_counters->manifold_count;
    return unresolved_status(*ctx->host_counters);
}

physgpu_status physgpu_epa_unresolved(physgpu_context* ctx, physgpu_pair* out, uint64_t capacity,
                                      uint64_t* count)
{
    if (!ctx || !count)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or count is NULL");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    physgpu_status st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    const uint32_t total = ctx->host_counters->epa_unresolved;
    *count = total;
    const uint64_t held = std::min<uint64_t>(total, ctx->epa_record_cap);
    if (out && held && capacity) {
        std::vector<uint64_t> tmp(held);
        PG_TRY_HIP(hipMemcpyAsync(tmp.data(), ctx->epa_unresolved_pairs, held * sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, ctx->stream));
        PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
        std::sort(tmp.begin(), tmp.end());
        for (uint64_t k = 0; k < std::min<uint64_t>(held, capacity); ++k) {
            out[k].first = uint32_t(tmp[k] >> 32);
            out[k].second = uint32_t(tmp[k]);
        }
    }
    if (held < total)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "more unresolved pairs than the context lists");
    if (out && capacity < held)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "output capacity below the unresolved count");
    return PHYSGPU_OK;
}

physgpu_status physgpu_narrowphase_stats_get(physgpu_context* ctx, physgpu_narrowphase_stats* s)
{
    if (!ctx || !s)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or stats is NULL");
    physgpu_status st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    fill_np_stats(*ctx->host_counters, s);
    return PHYSGPU_OK;
}

physgpu_status physgpu_step_stats_get(physgpu_context* ctx, physgpu_step_stats* s)
{
    if (!ctx || !s)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or stats is NULL");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    physgpu_status st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    collect_times(ctx);
    std::memset(s, 0, sizeof(*s));
    const Counters& c = *ctx->host_counters;
    st = fill_broad_stats(ctx, &s->broad_phase);
    if (st != PHYSGPU_OK)
        return st;
    fill_np_stats(c, &s->narrow_phase);
    if (ctx->timing) {
        s->broad_phase_ms = ctx->last_broadphase_ms;
        s->narrow_phase_ms = ctx->last_narrowphase_ms;
    }
    // possiblePairs (physics_world.cpp:571-577): each collider pairs with the
    // earlier ones of its world that sit on other bodies.
    const uint32_t n = ctx->n;
    std::vector<uint32_t> off;
    if (ctx->n_worlds)
        off = ctx->host_world_offsets;
    else
        off = {0u, n};
    std::unordered_map<uint32_t, uint64_t> per_body;
    uint64_t possible = 0;
    for (size_t w = 0; w + 1 < off.size(); ++w) {
        const uint64_t len = off[w + 1] - off[w];
        possible += len * (len - (len > 0 ? 1 : 0)) / 2;
        if (!ctx->host_body.empty() && ctx->has_world) {
            per_body.clear();
            for (uint32_t i = off[w]; i < off[w + 1]; ++i)
                possible -= per_body[ctx->host_body[i]]++;
        }
    }
    s->possible_pairs = possible;
    uint64_t candidates = 0;
    for (int q = 0; q < kClasses; ++q)
        candidates += c.class_count[q];
    s->candidate_pairs = candidates;
    s->contact_count = c.manifold_count;
    return PHYSGPU_OK;
}

}  // extern "C"

namespace pg {
namespace {

// Shared body of physgpu_contact_adjoint / physgpu_contact_geometry.
physgpu_status contact_pass(physgpu_context* ctx, uint64_t m, const float* g_normal,
                            const float* g_depth, const float* g_point, float* grad_pose,
                            float* grad_pair, float* geom_normal, float* geom_depth,
                            float* geom_point, physgpu_memory memory, uint64_t* skipped)
{
    if (memory != PHYSGPU_HOST && memory != PHYSGPU_DEVICE)
        return fail(PHYSGPU_INVALID_ARGUMENT, "unknown memory kind");
    if (!ctx->colliders_valid)
        return fail(PHYSGPU_INVALID_ARGUMENT, "no colliders set");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    physgpu_status st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    if (m != ctx->host_counters->manifold_count)
        return fail(PHYSGPU_INVALID_ARGUMENT, "m must equal the manifold count of the last step");
    const size_t M = size_t(m), N = size_t(ctx->n);
    const float* dn = g_normal;
    const float* dd = g_depth;
    const float* dp = g_point;
    float* gpose = grad_pose;
    float* gpair = grad_pair;
    float* on = geom_normal;
    float* od = geom_depth;
    float* op = geom_point;
    if (memory == PHYSGPU_HOST) {
        // Staging: inputs 19 m, pair gradients 14 m, pose gradients 7 n, geometry 19 m.
        const size_t need = 52 * std::max<size_t>(M, 1) + 7 * N;
        if (need > ctx->adjoint_capacity) {
            PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
            if (ctx->adjoint_scratch)
                hipFree(ctx->adjoint_scratch);
            ctx->adjoint_scratch = nullptr;
            ctx->adjoint_capacity = 0;
            PG_TRY_HIP(alloc(ctx->adjoint_scratch, need));
            ctx->adjoint_capacity = need;
        }
        float* s = ctx->adjoint_scratch;
        auto stage_in = [&](const float* src, size_t planes, float* dst) -> const float* {
            if (!src || !M)
                return nullptr;
            hipMemcpyAsync(dst, src, planes * M * 4, hipMemcpyHostToDevice, ctx->stream);
            return dst;
        };
        dn = stage_in(g_normal, 3, s);
        dd = stage_in(g_depth, 4, s + 3 * M);
        dp = stage_in(g_point, 12, s + 7 * M);
        gpair = grad_pair ? s + 19 * M : nullptr;
        gpose = grad_pose ? s + 33 * M : nullptr;
        on = geom_normal ? s + 33 * M + 7 * N : nullptr;
        od = geom_depth ? s + 36 * M + 7 * N : nullptr;
        op = geom_point ? s + 40 * M + 7 * N : nullptr;
    }
    launch_contact_adjoint(ctx, uint32_t(M), dn, dd, dp, on, od, op, gpair, gpose,
                           ctx->adjoint_skipped);
    PG_TRY_HIP(hipGetLastError());
    if (memory == PHYSGPU_HOST) {
        auto out = [&](float* dst, const float* src, size_t count) -> hipError_t {
            if (!dst || !count)
                return hipSuccess;
            return hipMemcpyAsync(dst, src, count * 4, hipMemcpyDeviceToHost, ctx->stream);
        };
        PG_TRY_HIP(out(grad_pair, gpair, 14 * M));
        PG_TRY_HIP(out(grad_pose, gpose, 7 * N));
        PG_TRY_HIP(out(geom_normal, on, 3 * M));
        PG_TRY_HIP(out(geom_depth, od, 4 * M));
        PG_TRY_HIP(out(geom_point, op, 12 * M));
    }
    uint32_t sk = 0;
    PG_TRY_HIP(hipMemcpyAsync(&sk, ctx->adjoint_skipped, 4, hipMemcpyDeviceToHost, ctx->stream));
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    if (skipped)
        *skipped = sk;
    return PHYSGPU_OK;
}

}  // namespace
}  // namespace pg

namespace pg {
namespace {

physgpu_status tree_alloc(physgpu_context* ctx)
{
    if (ctx->tree_tight)
        return PHYSGPU_OK;
    const size_t N = ctx->cap_n, P = ctx->cap_p;
    hipError_t e = hipSuccess;
    auto A = [&](hipError_t r) {
        if (e == hipSuccess)
            e = r;
    };
    A(alloc(ctx->tree_tight, 2 * N));
    A(alloc(ctx->tree_fat, 2 * N));
    A(alloc(ctx->tree_rank, N));
    A(alloc(ctx->tree_rank_user, N));
    A(alloc(ctx->tree_stage_ids, N));
    A(alloc(ctx->tree_stage_bounds, N));
    A(alloc(ctx->tree_stage_flags, N));
    A(alloc(ctx->tree_stage_pairs, P));
    if (e != hipSuccess) {
        void* ptrs[] = {ctx->tree_tight, ctx->tree_fat, ctx->tree_rank, ctx->tree_rank_user,
                        ctx->tree_stage_ids, ctx->tree_stage_bounds, ctx->tree_stage_flags,
                        ctx->tree_stage_pairs};
        for (void* p : ptrs)
            if (p)
                hipFree(p);
        ctx->tree_tight = ctx->tree_fat = nullptr;
        ctx->tree_rank = ctx->tree_stage_ids = nullptr;
        ctx->tree_rank_user = nullptr;
        ctx->tree_stage_bounds = nullptr;
        ctx->tree_stage_flags = nullptr;
        ctx->tree_stage_pairs = nullptr;
        return check_hip(e, "proxy tree allocation");
    }
    return PHYSGPU_OK;
}

// validateBounds, dynamic_aabb_tree.cpp:34-42.
bool finite_ordered(const physgpu_aabb& b)
{
    for (int a = 0; a < 3; ++a)
        if (!(std::isfinite(b.min[a]) && std::isfinite(b.max[a]) && b.min[a] <= b.max[a]))
            return false;
    return true;
}

bool tree_proxy_live(const physgpu_context* ctx, uint32_t id)
{
    return id < ctx->tree_live.size() && ctx->tree_live[id];
}

// Ids (and bounds) of one call into the device staging buffers.
physgpu_status tree_stage(physgpu_context* ctx, uint32_t count, const uint32_t* ids,
                          const physgpu_aabb* bounds)
{
    PG_TRY_HIP(hipMemcpyAsync(ctx->tree_stage_ids, ids, sizeof(uint32_t) * count,
                              hipMemcpyHostToDevice, ctx->stream));
    if (bounds)
        PG_TRY_HIP(hipMemcpyAsync(ctx->tree_stage_bounds, bounds, sizeof(physgpu_aabb) * count,
                                  hipMemcpyHostToDevice, ctx->stream));
    return PHYSGPU_OK;
}

}  // namespace
}  // namespace pg

extern "C" {

physgpu_status physgpu_contact_adjoint(physgpu_context* ctx, uint64_t m, const float* g_normal,
                                       const float* g_depth, const float* g_point,
                                       float* grad_pose, float* grad_pair, physgpu_memory memory,
                                       uint64_t* skipped)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    return contact_pass(ctx, m, g_normal, g_depth, g_point, grad_pose, grad_pair, nullptr,
                        nullptr, nullptr, memory, skipped);
}

physgpu_status physgpu_contact_geometry(physgpu_context* ctx, uint64_t m, float* normal,
                                        float* depth, float* point, physgpu_memory memory)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    return contact_pass(ctx, m, nullptr, nullptr, nullptr, nullptr, nullptr, normal, depth, point,
                        memory, nullptr);
}

physgpu_status physgpu_tree_create_proxies(physgpu_context* ctx, uint32_t count,
                                           const physgpu_aabb* bounds,
                                           const uint64_t* user_index, uint32_t* proxy_out)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    if (count == 0)
        return PHYSGPU_OK;
    if (!bounds || !user_index || !proxy_out)
        return fail(PHYSGPU_INVALID_ARGUMENT, "bounds, user_index and proxy_out are required");
    if (ctx->tree_user_proxy.size() + count > ctx->cap_n)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "more proxies than the context holds");
    std::unordered_map<uint64_t, uint32_t> batch;
    for (uint32_t i = 0; i < count; ++i) {
        if (!finite_ordered(bounds[i]))
            return fail(PHYSGPU_INVALID_ARGUMENT, "Dynamic AABB tree requires finite, ordered bounds");
        if (ctx->tree_user_proxy.count(user_index[i]) || !batch.emplace(user_index[i], i).second)
            return fail(PHYSGPU_INVALID_ARGUMENT, "user index already held by a live proxy");
    }
    PG_TRY_HIP(hipSetDevice(ctx->device));
    physgpu_status st = tree_alloc(ctx);
    if (st != PHYSGPU_OK)
        return st;
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t id;
        if (!ctx->tree_free.empty()) {
            id = ctx->tree_free.back();
            ctx->tree_free.pop_back();
        } else {
            id = uint32_t(ctx->tree_live.size());
            ctx->tree_live.push_back(0);
            ctx->tree_user.push_back(0);
        }
        ctx->tree_live[id] = 1;
        ctx->tree_user[id] = user_index[i];
        ctx->tree_user_proxy[user_index[i]] = id;
        proxy_out[i] = id;
    }
    st = tree_stage(ctx, count, proxy_out, bounds);
    if (st != PHYSGPU_OK)
        return st;
    launch_tree_create(ctx, count, ctx->tree_stage_ids, ctx->tree_stage_bounds);
    PG_TRY_HIP(hipGetLastError());
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    ctx->tree_created += count;
    ctx->tree_ranks_current = false;
    ctx->tree_build_current = false;
    return PHYSGPU_OK;
}

physgpu_status physgpu_tree_destroy_proxies(physgpu_context* ctx, uint32_t count,
                                            const uint32_t* proxy)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    if (count == 0)
        return PHYSGPU_OK;
    if (!proxy)
        return fail(PHYSGPU_INVALID_ARGUMENT, "proxy is NULL");
    std::unordered_map<uint32_t, uint32_t> seen;
    for (uint32_t i = 0; i < count; ++i)
        if (!tree_proxy_live(ctx, proxy[i]) || !seen.emplace(proxy[i], i).second)
            return fail(PHYSGPU_INVALID_ARGUMENT, "Invalid dynamic AABB tree proxy");
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t id = proxy[i];
        ctx->tree_live[id] = 0;
        ctx->tree_user_proxy.erase(ctx->tree_user[id]);
        ctx->tree_free.push_back(id);
    }
    ctx->tree_destroyed += count;
    ctx->tree_ranks_current = false;
    ctx->tree_build_current = false;
    return PHYSGPU_OK;
}

physgpu_status physgpu_tree_update_proxies(physgpu_context* ctx, uint32_t count,
                                           const uint32_t* proxy, const physgpu_aabb* bounds,
                                           uint8_t* reinserted)
{
    if (!ctx)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx is NULL");
    if (count == 0)
        return PHYSGPU_OK;
    if (!proxy || !bounds)
        return fail(PHYSGPU_INVALID_ARGUMENT, "proxy and bounds are required");
    std::unordered_map<uint32_t, uint32_t> seen;
    for (uint32_t i = 0; i < count; ++i) {
        if (!tree_proxy_live(ctx, proxy[i]))
            return fail(PHYSGPU_INVALID_ARGUMENT, "Invalid dynamic AABB tree proxy");
        if (!seen.emplace(proxy[i], i).second)
            return fail(PHYSGPU_INVALID_ARGUMENT, "a proxy may be updated once per call");
        if (!finite_ordered(bounds[i]))
            return fail(PHYSGPU_INVALID_ARGUMENT, "Dynamic AABB tree requires finite, ordered bounds");
    }
    PG_TRY_HIP(hipSetDevice(ctx->device));
    physgpu_status st = tree_stage(ctx, count, proxy, bounds);
    if (st != PHYSGPU_OK)
        return st;
    launch_tree_update(ctx, count, ctx->tree_stage_ids, ctx->tree_stage_bounds,
                       ctx->tree_stage_flags);
    PG_TRY_HIP(hipGetLastError());
    std::vector<uint8_t> flags(count);
    PG_TRY_HIP(hipMemcpyAsync(flags.data(), ctx->tree_stage_flags, count, hipMemcpyDeviceToHost,
                              ctx->stream));
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    for (uint32_t i = 0; i < count; ++i) {
        ctx->tree_reinserted += flags[i];
        if (reinserted)
            reinserted[i] = flags[i];
    }
    return PHYSGPU_OK;
}

physgpu_status physgpu_tree_size(physgpu_context* ctx, uint32_t* live)
{
    if (!ctx || !live)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or live is NULL");
    *live = uint32_t(ctx->tree_user_proxy.size());
    return PHYSGPU_OK;
}

physgpu_status physgpu_tree_find_pairs(physgpu_context* ctx, int policy, physgpu_proxy_pair* out,
                                       uint64_t capacity, uint64_t* count,
                                       physgpu_broadphase_stats* stats)
{
    if (!ctx || !count)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or count is NULL");
    if (policy != PHYSGPU_TREE_REBUILD && policy != PHYSGPU_TREE_REFIT)
        return fail(PHYSGPU_INVALID_ARGUMENT, "unknown tree policy");
    PG_TRY_HIP(hipSetDevice(ctx->device));
    physgpu_status st = tree_alloc(ctx);
    if (st != PHYSGPU_OK)
        return st;
    const uint32_t live = uint32_t(ctx->tree_user_proxy.size());
    const uint32_t slots = uint32_t(ctx->tree_live.size());
    if (!ctx->tree_ranks_current) {
        // Ranks of the live user indices (ascending), per proxy id.
        std::vector<std::pair<uint64_t, uint32_t>> order;
        order.reserve(live);
        for (uint32_t id = 0; id < slots; ++id)
            if (ctx->tree_live[id])
                order.emplace_back(ctx->tree_user[id], id);
        std::sort(order.begin(), order.end());
        std::vector<uint32_t> rank(slots, 0xFFFFFFFFu);
        std::vector<uint64_t> user(live);
        for (uint32_t r = 0; r < live; ++r) {
            rank[order[r].second] = r;
            user[r] = order[r].first;
        }
        if (slots)
            PG_TRY_HIP(hipMemcpyAsync(ctx->tree_rank, rank.data(), sizeof(uint32_t) * slots,
                                      hipMemcpyHostToDevice, ctx->stream));
        if (live)
            PG_TRY_HIP(hipMemcpyAsync(ctx->tree_rank_user, user.data(), sizeof(uint64_t) * live,
                                      hipMemcpyHostToDevice, ctx->stream));
        PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
        ctx->tree_ranks_current = true;
    }
    const uint32_t resident_n = ctx->n;
    const bool refit = policy == PHYSGPU_TREE_REFIT && ctx->tree_build_current && live >= 2;
    ctx->n = live;
    launch_step_begin(ctx);
    launch_tree_records(ctx, slots);
    if (live >= 2) {
        if (refit) {
            launch_lbvh_refit(ctx);
        } else {
            launch_aabb_reductions(ctx);
            launch_lbvh(ctx);
        }
    }
    launch_canonical_pairs(ctx);
    launch_tree_pairs_out(ctx, ctx->tree_stage_pairs);
    ctx->n = resident_n;
    ctx->resident_aabbs_valid = false;
    ctx->tree_pairs_loaded = true;
    ctx->tree_build_current = live >= 2;
    PG_TRY_HIP(hipGetLastError());
    st = read_counters(ctx);
    if (st != PHYSGPU_OK)
        return st;
    const Counters& c = *ctx->host_counters;
    if (live < 2)
        ctx->tree_height = 0;
    else if (!refit)
        ctx->tree_height = c.tree_height;
    *count = c.pair_count;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->aabb_pairs = c.pair_count;
        stats->tree_proxy_count = live;
        stats->tree_node_pair_visits = c.tree_visits;
        stats->tree_leaf_checks = c.tree_leaf_checks;
        stats->tree_height = ctx->tree_height;
        stats->tree_insertions = ctx->tree_created;
        stats->tree_removals = ctx->tree_destroyed;
        stats->tree_reinsertions = ctx->tree_reinserted;
    }
    ctx->tree_created = ctx->tree_destroyed = ctx->tree_reinserted = 0;
    if (c.tree_overflow)
        return fail(PHYSGPU_DEVICE_ERROR, "LBVH traversal stack overflow");
    if (c.pair_count > ctx->cap_p)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "candidate pairs exceed the context capacity");
    const uint64_t n_out = std::min<uint64_t>(capacity, c.pair_count);
    if (out && n_out) {
        PG_TRY_HIP(hipMemcpyAsync(out, ctx->tree_stage_pairs, sizeof(physgpu_proxy_pair) * n_out,
                                  hipMemcpyDeviceToHost, ctx->stream));
        PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (out && capacity < c.pair_count)
        return fail(PHYSGPU_CAPACITY_EXCEEDED, "output capacity below the pair count");
    return PHYSGPU_OK;
}

physgpu_status physgpu_kernel_times(physgpu_context* ctx, const char** names, double* avg_ms,
                                    uint64_t* launches, uint32_t capacity, uint32_t* count)
{
    if (!ctx || !count)
        return fail(PHYSGPU_INVALID_ARGUMENT, "ctx or count is NULL");
    PG_TRY_HIP(hipStreamSynchronize(ctx->stream));
    collect_times(ctx);
    *count = uint32_t(ctx->kernel_times.size());
    for (uint32_t i = 0; i < capacity && i < ctx->kernel_times.size(); ++i) {
        const KernelTime& k = ctx->kernel_times[i];
        if (names)
            names[i] = k.name;
        if (avg_ms)
            avg_ms[i] = k.launches ? k.total_ms / double(k.launches) : 0.0;
        if (launches)
            launches[i] = k.launches;
    }
    return PHYSGPU_OK;
}

}  // extern "C"

// ---- Allocation-free broad phase (physgpu_find_candidate_pairs_noalloc) ----
// The working memory of one broad phase of at most PHYSGPU_NOALLOC_MAX_BODIES
// bodies and PHYSGPU_NOALLOC_MAX_PAIRS pairs, carved from one fixed array per
// device (the library holds it on every device it runs on until the process
// ends).  Host code never passes `scratch` itself (a host-side stand-in):
// it uses the per-device address from hipGetSymbolAddress.

namespace pg {
namespace {

constexpr uint32_t kNoAllocBodies = PHYSGPU_NOALLOC_MAX_BODIES;
constexpr uint64_t kNoAllocPairs = PHYSGPU_NOALLOC_MAX_PAIRS;
constexpr size_t kScratchAlign = 256;

uint32_t noalloc_table_bits()
{
    return grid_table_bits(kNoAllocBodies);
}

// Assigns the broad-phase buffers of `c` (sized for the maximum) from `base`
// in the order of physgpu_create; returns the bytes used.  With base == NULL
// only the size is computed.
size_t carve_noalloc(physgpu_context& c, char* base)
{
    size_t off = 0;
    auto take = [&](auto*& p, size_t count) {
        using T = std::remove_reference_t<decltype(*p)>;
        off = (off + kScratchAlign - 1) & ~(kScratchAlign - 1);
        p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += std::max<size_t>(count, 1) * sizeof(T);
    };
    const size_t N = kNoAllocBodies, P = kNoAllocPairs;
    const size_t tiles = std::max<size_t>(div_up(N, kSortMinTile), div_up(P, kSortMinTile));
    take(c.aabb, 2 * N);
    take(c.sorted_aabb, 2 * N);
    take(c.items_a, N);
    take(c.items_b, N);
    take(c.histogram, 256 * tiles);
    take(c.digit_total, 256);
    take(c.cell_table, 2 * (size_t(1) << noalloc_table_bits()));
    take(c.overflow_list, N);
    take(c.grid_over_ctr, 1);
    take(c.bucket_bounds, 2 * kMaxGridBuckets);
    take(c.node_left, N);
    take(c.node_right, N);
    take(c.node_parent, 2 * N);
    take(c.node_last, N);
    take(c.node_box, 2 * N);
    take(c.node_first, N);
    take(c.level_box, 2 * lbvh_level_boxes(N));
    take(c.pairs_raw, P);
    take(c.pairs_tmp, P);
    take(c.seg_off, N + 1);
    take(c.seg_tiles, div_up(N + 1, 2048));
    take(c.pair_rank, P);
    take(c.pair_perm, P);
    take(c.sort_scratch, P);
    take(c.sort_scratch_perm, P);
    take(c.seg_lists, 2 * N);
    take(c.partials, size_t(kPartialSlots) * kMaxPartialBlocks);
    take(c.counters, 1);
    c.hist_capacity = uint32_t(256 * tiles);
    return off;
}

// Upper bound of carve_noalloc's size (each buffer plus its alignment).
constexpr size_t noalloc_bytes()
{
    const size_t N = kNoAllocBodies, P = kNoAllocPairs, a = kScratchAlign;
    const size_t tiles = (std::max(N, size_t(P)) + kSortMinTile - 1) / kSortMinTile;
    return 28 * a + 2 * N * 16 * 2 + N * 8 * 2 + 256 * tiles * 4 + 256 * 4 + 4 - N * 4 +
           2 * kMaxGridBuckets * 4 +
           2 * (size_t(1) << grid_table_bits(N)) * 4 + N * 4 * 2 + N * 4 * 2 + 2 * N * 4 + N * 4 +
           2 * N * 16 + N * 4 + 2 * lbvh_level_boxes(N) * 16 + P * 8 * 2 +
           (N + 1) * 4 + ((N + 1 + 2047) / 2048) * 4 + P * 4 * 2 + P * 8 + P * 4 + 2 * N * 4 +
           size_t(kPartialSlots) * kMaxPartialBlocks * 8 + sizeof(Counters);
}

}  // namespace
}  // namespace pg

// The fixed working memory (one copy per device).
static __device__ __align__(256) unsigned char scratch[pg::noalloc_bytes()];

namespace pg {
namespace {

constexpr int kMaxNoAllocDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxNoAllocDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current).
physgpu_status scratch_on(int device, void** out)
{
    if (device < 0 || device >= kMaxNoAllocDevices)
        return fail(PHYSGPU_UNSUPPORTED, "device index beyond the allocation-free table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return check_hip(e, "hipGetSymbolAddress(scratch)");
        if (!p)
            return fail(PHYSGPU_DEVICE_ERROR, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return PHYSGPU_OK;
}

// Canonical pairs (first << 32 | second) to physgpu_pair records, at most
// `capacity`, and the full pair count (~0 when the LBVH traversal stack
// overflowed).
__global__ __launch_bounds__(kBlock) void k_noalloc_out(const uint64_t* __restrict__ pairs,
                                                        const Counters* __restrict__ c,
                                                        physgpu_pair* __restrict__ out,
                                                        uint64_t capacity,
                                                        uint64_t* __restrict__ count)
{
    const uint64_t total = c->pair_count;
    const uint64_t held = min(min(total, uint64_t(c->sort_count)), capacity);
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < held; i += stride) {
        const uint64_t p = pairs[i];
        out[i] = physgpu_pair{uint32_t(p >> 32), uint32_t(p)};
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *count = c->tree_overflow ? ~0ull : total;
}

}  // namespace
}  // namespace pg

extern "C" {

physgpu_status physgpu_find_candidate_pairs_noalloc(const physgpu_aabb* bounds, uint32_t n,
                                                    int algorithm, double grid_cell_size,
                                                    physgpu_pair* out, uint64_t capacity,
                                                    uint64_t* count, void* hip_stream)
{
    if (!count || (n && !bounds) || (capacity && !out))
        return fail(PHYSGPU_INVALID_ARGUMENT, "count is NULL, or bounds / out missing");
    if (algorithm < PHYSGPU_SWEEP_AND_PRUNE || algorithm > PHYSGPU_DYNAMIC_TREE)
        return fail(PHYSGPU_INVALID_ARGUMENT, "unknown broadphase algorithm");
    if (n > kNoAllocBodies)
        return fail(PHYSGPU_INVALID_ARGUMENT,
                    "n = " + std::to_string(n) + " is above PHYSGPU_NOALLOC_MAX_BODIES (" +
                        std::to_string(kNoAllocBodies) + ")");
    int device = 0;
    PG_TRY_HIP(hipGetDevice(&device));
    void* base = nullptr;
    physgpu_status st = scratch_on(device, &base);
    if (st != PHYSGPU_OK)
        return st;
    // A context on the stack over the fixed memory (nothing is allocated).
    physgpu_context c;
    c.device = device;
    c.tuning = default_tuning();
    c.stream = static_cast<hipStream_t>(hip_stream);
    c.cap_n = kNoAllocBodies;
    c.cap_p = kNoAllocPairs;
    c.table_bits = noalloc_table_bits();
    if (carve_noalloc(c, static_cast<char*>(base)) > sizeof(scratch))
        return fail(PHYSGPU_DEVICE_ERROR, "allocation-free layout exceeds its fixed memory");
    c.n = n;
    c.pairs = c.pairs_raw;
    launch_step_begin(&c);
    if (n)
        launch_convert_aabbs_in(&c, bounds);
    run_broadphase(&c, algorithm, grid_cell_size, false);
    hipLaunchKernelGGL(k_noalloc_out, dim3(256), dim3(kBlock), 0, c.stream, c.pairs, c.counters, out,
                       capacity, count);
    PG_TRY_HIP(hipGetLastError());
    return PHYSGPU_OK;
}

}  // extern "C"
