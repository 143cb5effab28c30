// Internal context of the collision library: device buffers (SoA planes in
// HBM), device counters, launch and timing helpers.  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/physgpu.h"

namespace pg {

constexpr int kClasses = 6;        // physgpu_shape_pair values
constexpr uint32_t kBlock = 256;   // threads per workgroup for streaming kernels
constexpr uint32_t kStrideGrid = 2048;  // workgroups of grid-stride kernels (8 per CU)

// Radix sort tiling: 256 threads x (4, 8 or 16) items per workgroup (the
// tuning's sort_items); buffers are sized for the smallest tile.
constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortMinTile = kSortThreads * 4;
// LBVH level boxes: aligned blocks of 4^k leaves, k = 1 .. at most this.
constexpr uint32_t kLbvhMaxLevels = 16;
// Level boxes for n leaves: sum of ceil(n / 4^k) <= n / 3 + kLbvhMaxLevels.
constexpr size_t lbvh_level_boxes(size_t n) { return n / 3 + kLbvhMaxLevels + 1; }

// Compaction tiling.
constexpr uint32_t kScanTile = 256;  // candidates per compaction tile

// Per-workgroup partial results.  Device-scope atomics on one address are
// serialised (about 11 ns each on MI355X), so statistics and reductions are
// written as one partial per workgroup into a slot of `partials` and folded
// by k_fold (one small launch) instead.  Kernels that write partials are
// launched with at most kMaxPartialBlocks workgroups.
constexpr uint32_t kMaxPartialBlocks = 8192;
// EPA workgroups (one wavefront each) at most.
constexpr uint32_t kMaxEpaGroups = 1u << 20;
// EPA runs past the LDS polytope bounds: 32-bit words of one hand-over
// record (iteration, vertices, faces), workgroups of the continuation kernel.
constexpr uint32_t kEpaRecordWords = 1536;
constexpr uint32_t kEpaContinueGroups = 256;
// Workgroups of the consumer that continues handed-over runs while GJK + EPA
// still runs, and how long it waits for them at most (s_memrealtime ticks of
// 10 ns).
constexpr uint32_t kEpaEarlyGroups = 64;
constexpr uint64_t kEpaEarlyPollTicks = 500000;
constexpr uint64_t kEpaEarlyStartTicks = 5000;
enum PartialSlot : uint32_t {
    kSlotGridTests = 0,
    kSlotGridOverflowTests,
    kSlotWindowTests,
    kSlotTreeVisits,
    kSlotTreeLeaves,
    kSlotWidthMaxX,   // AABB reductions: max / max / sum / min x3 / max x3
    kSlotWidthMax,
    kSlotWidthSum,
    kSlotCentreMin,   // 3 slots
    kSlotCentreMax = kSlotCentreMin + 3,  // 3 slots
    kSlotContacts = kSlotCentreMax + 3,   // one per shape-pair class (6 slots)
    kSlotGjkFailures = kSlotContacts + 6,
    kSlotEpaOverflows,
    kSlotEpaItemsDone,  // GJK + EPA work items finished (the hand-over consumer's stop test)
    kPartialSlots
};
// EPA's partial slots: contacts of classes 3..5, failures, overflows, items
// done, written by atomic adds spread over kEpaPartials addresses each and
// zeroed before each narrow phase (by the classification, or by a grid
// step's key pass).
constexpr uint32_t kEpaPartialFirst = kSlotContacts + 3;
constexpr uint32_t kEpaPartialSlots = 6;
constexpr uint32_t kEpaPartials = 1024;
enum FoldOp : uint32_t { kFoldSum64 = 0, kFoldSum32, kFoldMax32, kFoldMin32 };
struct FoldEntry {
    uint32_t slot;    // PartialSlot
    uint32_t blocks;  // workgroups that wrote the slot
    uint32_t op;      // FoldOp
    uint32_t offset;  // byte offset of the destination in Counters
};
constexpr uint32_t kMaxFold = 12;
struct FoldList {
    FoldEntry e[kMaxFold];
    uint32_t count;
    uint32_t finish_pairs;  // also clamp the pair count to finish_cap (see launch_finish_pairs)
    uint64_t finish_cap;
};

// Device-side counters of one collision step.  Zeroed by the step-begin
// kernel; read back by the host only when it asks for counts.
struct Counters {
    unsigned long long pair_count;      // pairs emitted by the broad phase (may exceed capacity)
    unsigned long long window_tests;    // SAP x-window comparisons
    unsigned long long grid_tests;      // grid AABB tests (incl. overflow tests)
    unsigned long long tree_visits;     // LBVH node visits
    unsigned long long tree_leaf_checks;
    unsigned long long grid_overflow_tests;  // overflow-path AABB tests (added to grid_tests)
    unsigned long long width_sum_fx;    // sum of max AABB sides, 2^-16 fixed point
    uint32_t grid_entries;              // bodies placed in a cell
    uint32_t overflow_count;            // grid bodies tested against everything
    uint32_t wmax_x_key;                // float_key of the largest x side
    uint32_t wmax_key;                  // float_key of the largest side over all axes
    uint32_t cmin_key[3];               // float_key bounds of AABB centres
    uint32_t cmax_key[3];
    uint32_t class_count[kClasses];     // candidates per shape-pair class
    uint32_t class_contacts[kClasses];  // contacts per class
    uint32_t gjk_failures;              // GJK hit, EPA did not converge
    uint32_t epa_overflow;              // EPA exceeded its bounded polytope
    uint32_t manifold_count;
    uint32_t tree_height;
    uint32_t sort_count;                // min(pair_count, capacity): pairs held
    uint32_t cell_bits;                 // grid cell size used (float bits)
    uint32_t tree_overflow;             // LBVH traversal stack overflows (must stay 0)
    uint32_t mid_segments;              // pair segments sorted by one wave
    uint32_t long_segments;             // pair segments sorted by one workgroup
    uint32_t epa_continued;             // EPA runs handed over past the LDS bounds
    uint32_t epa_unresolved;            // hull pairs left without an EPA result
};

struct Timer {
    const char* name;
    hipEvent_t start, stop;
    hipStream_t stream;
};

struct KernelTime {
    const char* name;
    double total_ms;
    uint64_t launches;
};

}  // namespace pg

struct physgpu_context {
    int device = 0;
    physgpu_tuning tuning{};
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    // Second stream of the narrow phase: the SS / SB / BB kernels run on it
    // while the main stream runs EPA (fork and join by events, so the step
    // stays one capturable sequence on the main stream).  This and the third
    // stream are the device's, shared by all its contexts.
    hipStream_t side_stream = nullptr;
    hipEvent_t sort_event = nullptr;  // pairs emitted: the canonical sort may start
    hipEvent_t fork_event = nullptr;
    hipEvent_t join_event = nullptr;
    // Third stream: the consumer of EPA runs handed over past the LDS bounds
    // (k_epa_early) runs on it beside GJK + EPA; joined before compaction.
    hipStream_t early_stream = nullptr;
    hipEvent_t early_join_event = nullptr;
    uint32_t cap_n = 0;
    uint64_t cap_p = 0;
    uint32_t cap_h = 0;

    // Colliders: the uploaded planes (pose, shape_type, shape_data,
    // hull_range) and the resident layouts every kernel reads, 2 float4 per
    // body each (pose4: load_pose4; shape4: shape_data / shape_type /
    // hull_range), rebuilt by launch_pose_aos / launch_shape_aos on upload.
    uint32_t n = 0;
    uint32_t n_hull_points = 0;
    float* pose = nullptr;        // 7 x cap_n
    uint8_t* shape_type = nullptr;
    float* shape_data = nullptr;  // 3 x cap_n
    uint32_t* hull_range = nullptr;  // 2 x cap_n
    float4* pose4 = nullptr;         // 2 x cap_n
    float4* shape4 = nullptr;        // 2 x cap_n
    float* hull_points = nullptr;    // 3 x cap_h
    bool has_hulls = false;
    // Largest AABB side bound of the shapes (max key, sum) and the collision
    // step's grid cell from it (float bits in word 1), per upload.
    uint32_t* shape_cell = nullptr;
    uint32_t* hull_list = nullptr;        // hull collider indices (cap_n), built per upload
    uint32_t* hull_list_count = nullptr;  // their count (device)
    bool colliders_valid = false;

    // AABB records (2 float4 per body, see wave.hpp Rec) and their sorted copy
    // (the record's index field holds the original collider index).
    float4* aabb = nullptr;
    float4* sorted_aabb = nullptr;

    // Radix sort scratch (u64 items).
    uint64_t* items_a = nullptr;
    uint64_t* items_b = nullptr;
    uint32_t* histogram = nullptr;  // 256 x tiles
    uint32_t* digit_total = nullptr;  // 256 per pass
    uint32_t hist_capacity = 0;

    // Grid.
    uint32_t* cell_table = nullptr;  // 2 x table_size (start, end)
    uint32_t table_bits = 0;
    uint32_t* overflow_list = nullptr;
    // Slots taken in overflow_list by the key kernel (device; zero between
    // steps: the bucket kernel publishes it as the overflow count and zeroes it).
    uint32_t* grid_over_ctr = nullptr;
    // Per-bucket (start, end) after two radix passes (2 x kMaxGridBuckets),
    // zeroed by the key pass of each grid step.
    uint32_t* bucket_bounds = nullptr;

    // LBVH: internal nodes [0, n-1), leaves by sorted position.
    int32_t* node_left = nullptr;
    int32_t* node_right = nullptr;    // per split position s: the right child of the node
                                      // splitting s | s + 1 (the rope of subtrees ending at s)
    int32_t* node_parent = nullptr;   // internal parents (n-1) then leaf parents (n)
    uint32_t* node_first = nullptr;   // first leaf position covered by an internal node
    uint32_t* node_last = nullptr;    // last leaf position covered by an internal node
    float4* node_box = nullptr;       // internal node records: bounds, left child, rope
    float4* level_box = nullptr;      // boxes of aligned 4^k-leaf blocks, k = 1, 2, ...
    uint64_t* lbvh_sorted = nullptr;  // leaf order of the last build (items_a or items_b)

    // Pairs as first << 32 | second: raw emission buffer, sort scratch, and
    // the canonical (sorted) list, which aliases one of the two.
    uint64_t* pairs_raw = nullptr;
    uint64_t* pairs_tmp = nullptr;
    uint64_t* pairs = nullptr;
    // Canonical-order scratch: per-body segment offsets (cap_n + 1), their
    // scan tile sums, each pair's rank in its segment (cap_p), and the lists
    // of segments too long for one thread (2 x cap_n).
    uint32_t* seg_off = nullptr;
    uint32_t* seg_tiles = nullptr;
    uint32_t* pair_perm = nullptr;          // emission index of each sorted pair (cap_p)
    uint64_t* sort_scratch = nullptr;       // long-segment staging (cap_p)
    uint32_t* sort_scratch_perm = nullptr;  // (cap_p)
    uint32_t* pair_rank = nullptr;
    uint32_t* seg_lists = nullptr;

    // Persistent proxies (physgpu_tree_*), allocated on first use.  Device:
    // tight and fat bounds per proxy id (2 float4 each, record layout), the
    // rank of each live proxy's user index among the live ones (~0u when
    // dead), the user index of each rank, and host-to-device staging.
    float4* tree_tight = nullptr;
    float4* tree_fat = nullptr;
    uint32_t* tree_rank = nullptr;
    uint64_t* tree_rank_user = nullptr;
    uint32_t* tree_stage_ids = nullptr;
    physgpu_aabb* tree_stage_bounds = nullptr;
    uint8_t* tree_stage_flags = nullptr;
    physgpu_proxy_pair* tree_stage_pairs = nullptr;
    // Host bookkeeping: liveness and user index per proxy id, free ids
    // (reused last-in first-out), live user indices, counts since the last
    // query, and whether the ranks or the LBVH build are current.
    std::vector<uint8_t> tree_live;
    std::vector<uint64_t> tree_user;
    std::vector<uint32_t> tree_free;
    std::unordered_map<uint64_t, uint32_t> tree_user_proxy;
    uint64_t tree_created = 0, tree_destroyed = 0, tree_reinserted = 0;
    uint32_t tree_height = 0;
    bool tree_ranks_current = false;
    bool tree_build_current = false;  // last broad phase was this tree's LBVH build
    bool resident_aabbs_valid = true; // ctx->aabb holds the resident colliders' bounds
    bool finish_pending = false;      // the next fold also finishes the broad phase's pairs
    bool tree_pairs_loaded = false;   // the candidate pairs are a tree query's (rank space)
    bool np_classified = false;       // the broad phase has classified the pairs (grid step)

    // Narrow phase.  Its input pairs and, when they are not in canonical
    // order, the emission index of each canonical position (np_perm; null
    // means the input order is the output order).
    const uint64_t* np_pairs = nullptr;
    const uint32_t* np_perm = nullptr;
    uint32_t* buckets = nullptr;       // kClasses x cap_p candidate indices
    physgpu_manifold* slots = nullptr; // one record per candidate
    uint32_t* hit = nullptr;           // per candidate
    uint32_t* tile_sums = nullptr;     // compaction
    physgpu_manifold* manifolds = nullptr;  // compacted, candidate order
    physgpu_narrowphase_options np_options{};
    // EPA runs past the LDS polytope bounds: hand-over records
    // (kEpaRecordWords each) and the pairs left unresolved (first << 32 |
    // second), epa_record_cap of each.
    uint32_t* epa_records = nullptr;
    uint64_t* epa_unresolved_pairs = nullptr;
    uint32_t epa_record_cap = 0;
    // Per record: ready and claim words (epoch-tagged, 2 x epa_record_cap), then
    // "k_epa_continue has started" and "k_gjk_epa has started" (zeroed once
    // at create), and the epoch (one per narrow phase).
    uint32_t* epa_flags = nullptr;
    uint32_t* epa_epoch = nullptr;

    // World-level extras (optional): per collider, the pose of its body, its
    // body index, its local transform and its body's materials.
    float* world_body_pose = nullptr;  // 7 x cap_n
    uint32_t* body = nullptr;
    float* local_pose = nullptr;
    float* friction = nullptr;
    float* restitution = nullptr;
    bool has_world = false;
    bool world_has_local = false;
    bool world_has_body = false;
    bool world_has_friction = false;
    bool world_has_restitution = false;

    // Batch of worlds (n_worlds = 0: one world): collider offsets (n_worlds + 1),
    // world of every collider, per-world pair and manifold ranges of the last step.
    uint32_t n_worlds = 0;
    uint32_t world_capacity = 0;
    uint32_t* world_offsets = nullptr;
    uint32_t* world_of = nullptr;
    uint64_t* world_pair_off = nullptr;
    uint64_t* world_manifold_off = nullptr;

    // The reference grid's statistics (gridstats.hip), allocated on demand and
    // grown only: two scalar blocks (bytes), per-record cell ranges, the cell
    // table (tags, owners, counts) and two collision lists.
    unsigned char* gs_scalars = nullptr;
    int4* gs_ranges = nullptr;
    uint64_t* gs_tags = nullptr;
    uint64_t* gs_owner = nullptr;
    uint32_t* gs_counts = nullptr;
    uint64_t* gs_list = nullptr;
    size_t gs_scalars_cap = 0, gs_ranges_cap = 0, gs_tags_cap = 0, gs_owner_cap = 0,
           gs_counts_cap = 0, gs_list_cap = 0;

    // Adjoint staging for host-memory calls (floats) and its skip counter.
    float* adjoint_scratch = nullptr;
    size_t adjoint_capacity = 0;
    uint32_t* adjoint_skipped = nullptr;

    // HIP graph of the device-resident step and what it was captured with.
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_algorithm = -1;
    double graph_cell = 0.0;
    physgpu_narrowphase_options graph_options{};

    // Per-workgroup partials (kPartialSlots x kMaxPartialBlocks).
    uint64_t* partials = nullptr;

    // Counters (device) and a pinned host mirror.
    pg::Counters* counters = nullptr;
    pg::Counters* host_counters = nullptr;

    // Last step.
    int last_algorithm = -1;
    double last_cell_size = 0.0;
    // Stage times of the last timed step (ms; physgpu_step_stats).
    double last_broadphase_ms = 0.0;
    double last_narrowphase_ms = 0.0;
    // Host copies behind StepStats::possiblePairs: owning body per collider
    // (empty: one body per collider) and world offsets (empty: one world).
    std::vector<uint32_t> host_body;
    std::vector<uint32_t> host_world_offsets;

    // Timing.
    bool timing = false;
    std::vector<pg::Timer> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<pg::KernelTime> kernel_times;
};

namespace pg {

// Error helpers (thread-local message behind physgpu_last_error).
void set_error(const std::string& msg);
physgpu_status fail(physgpu_status st, const std::string& msg);
physgpu_status check_hip(hipError_t e, const char* what);

#define PG_TRY_HIP(expr)                                               \
    do {                                                               \
        hipError_t e_ = (expr);                                        \
        if (e_ != hipSuccess)                                          \
            return ::pg::check_hip(e_, #expr);                         \
    } while (0)

// Timing brackets around one launch (no-ops unless ctx->timing).
// Kernel timers (when timing is on): events on `stream`, the main stream by default.
void time_begin(physgpu_context* ctx, const char* name, Timer& t, hipStream_t stream = nullptr);
// Names of the stage brackets of a device-resident step.
extern const char* const kStageBroadphase;
extern const char* const kStageNarrowphase;
void time_end(physgpu_context* ctx, Timer& t);
void collect_times(physgpu_context* ctx);

inline uint32_t div_up(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }
inline uint32_t stride_grid(uint64_t work)
{
    uint64_t g = (work + kBlock - 1) / kBlock;
    if (g < 1)
        g = 1;
    return uint32_t(g < kStrideGrid ? g : kStrideGrid);
}

// ---- launchers implemented in the kernel translation units -------------
// aabb.hip
void launch_step_begin(physgpu_context* ctx);
// With with_stats, also the step counters' reset (launch_step_begin) and the
// width statistics behind the automatic grid cell (what
// launch_aabb_reductions computes), in the same launch.
void launch_compute_aabbs(physgpu_context* ctx, bool with_stats = false);
// Pose planes (ctx->pose, or a caller's device planes `src`) -> ctx->pose4.
void launch_pose_aos(physgpu_context* ctx, const float* src = nullptr);
void launch_shape_aos(physgpu_context* ctx);  // shape planes -> ctx->shape4
void launch_hull_list(physgpu_context* ctx);
// Grid cell of the collision step from the resident shapes (ctx->shape_cell),
// once per upload.
void launch_shape_cell(physgpu_context* ctx);
void launch_aabb_reductions(physgpu_context* ctx);
void launch_convert_aabbs_out(physgpu_context* ctx, physgpu_aabb* dev_out);
void launch_convert_aabbs_in(physgpu_context* ctx, const physgpu_aabb* dev_in);
void launch_compose_world(physgpu_context* ctx);
// radix_sort.hip: stable sort of n u64 items by `bits` bits starting at bit
// `shift0` (and, when bits1 > 0, then by `bits1` bits at `shift1`, which is the
// more significant field).  Result ends in *result (one of a, b).  With
// gather_dst, the final pass also writes gather_dst[p] = gather_src[index]
// (2 float4 records) for the item at sorted position p, index = item >> 32.
// With first_hist_done the caller has written the first pass's per-tile
// histogram (ctx->histogram[digit * tiles + tile], tiles of
// kSortThreads x ctx->tuning.sort_items items, n items on the host).
// Items with a bit of `drop` set are left out by the first pass (and must
// not be counted in a histogram the caller wrote): the result holds the
// others, in order, at the front; their count is written to *kept (device),
// which the later passes read as their item count.
void radix_sort(physgpu_context* ctx, uint64_t* a, uint64_t* b, uint32_t n,
                const uint32_t* device_count, uint32_t shift0, uint32_t bits0, uint32_t shift1,
                uint32_t bits1, uint64_t** result, const float4* gather_src = nullptr,
                float4* gather_dst = nullptr, bool first_hist_done = false,
                uint64_t drop = 0, uint32_t* kept = nullptr);
// sap.hip
void launch_sap_keys(physgpu_context* ctx);
void launch_gather_sorted(physgpu_context* ctx, const uint64_t* sorted_items);
void launch_sap_sweep(physgpu_context* ctx);
void launch_sap_window_tests(physgpu_context* ctx);
// grid.hip
// Key bits of the grid for n bodies in a context whose table has 2^table_bits
// entries: t table bits (3k with 2^3k >= 2n, k >= 2, t <= table_bits), Morton
// bits per axis k, bucket bits (radix-sorted, one pass up to 8, two up to
// 16) and the low bits counted per bucket (at most 13).
struct GridBits {
    uint32_t table, axis, bucket, low;
};
// Table bits of a context for up to `bodies` colliders: 3k with 2^3k >= 2
// bodies, 2 <= k <= 9.
constexpr uint32_t grid_table_bits(uint64_t bodies)
{
    uint32_t k = 2;
    while (k < 9 && (uint64_t(1) << (3 * k)) < 2 * bodies)
        ++k;
    return 3 * k;
}
GridBits grid_bits(uint32_t n, uint32_t table_bits);
constexpr uint32_t kMaxGridBuckets = 1u << 12;  // 27 table bits, 15 counted per bucket
void launch_grid(physgpu_context* ctx, double cell_size, bool fused = false);
// aabb.hip: AABB records of the resident colliders with the grid's keys.
void launch_aabbs_grid_keys(physgpu_context* ctx, float cell, const GridBits& g);
// gridstats.hip: the reference grid's counters (cell size, entries,
// comparisons, overflow AABBs) of the resident AABBs, per world of a batch,
// summed; runs on the main stream and waits for it.
physgpu_status reference_grid_stats(physgpu_context* ctx, physgpu_broadphase_stats* s);
// lbvh.hip
void launch_lbvh(physgpu_context* ctx);
void launch_lbvh_refit(physgpu_context* ctx);
void launch_lbvh_boxes_and_query(physgpu_context* ctx);
// tree.hip
void launch_tree_create(physgpu_context* ctx, uint32_t count, const uint32_t* ids,
                        const physgpu_aabb* bounds);
void launch_tree_update(physgpu_context* ctx, uint32_t count, const uint32_t* ids,
                        const physgpu_aabb* bounds, uint8_t* reinserted);
void launch_tree_records(physgpu_context* ctx, uint32_t slots);
void launch_tree_pairs_out(physgpu_context* ctx, physgpu_proxy_pair* out);
// pairs.hip
void launch_canonical_pairs(physgpu_context* ctx);  // finish + sort on the main stream
void launch_finish_pairs(physgpu_context* ctx);     // pair count clamp (main stream)
void launch_sort_pairs(physgpu_context* ctx, hipStream_t stream);
void launch_pairs_out(physgpu_context* ctx, physgpu_pair* dev_out, uint64_t count);
void launch_pairs_in(physgpu_context* ctx, const physgpu_pair* dev_in, uint64_t count);
// narrowphase.hip
void launch_np_begin(physgpu_context* ctx);
// Narrow phase over ctx->np_pairs; sort_on_side also builds the canonical
// pair order (and np_perm) on the side stream meanwhile.
void launch_narrowphase(physgpu_context* ctx, bool world, bool sort_on_side);
void launch_hull_adjoint(physgpu_context* ctx, uint32_t m, const float* g_normal,
                         const float* g_depth, const float* g_point, float* geom_normal,
                         float* geom_depth, float* geom_point, float* grad_pair, float* grad_pose,
                         uint32_t* skipped);
// adjoint.hip
void launch_contact_adjoint(physgpu_context* ctx, uint32_t m, const float* g_normal,
                            const float* g_depth, const float* g_point, float* geom_normal,
                            float* geom_depth, float* geom_point, float* grad_pair,
                            float* grad_pose, uint32_t* skipped);
// counters.hip: folds partial slots into the counters (entries with distinct
// destinations).
void launch_fold(physgpu_context* ctx, const FoldList& list);
// batch.hip
void launch_world_of(physgpu_context* ctx);
void launch_world_ranges(physgpu_context* ctx);

}  // namespace pg
