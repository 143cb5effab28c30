// Linear BVH standing in for DynamicAabbTree (src/collision/dynamic_aabb_tree.cpp).
//
// The reference tree's query returns every pair of proxies whose *tight*
// bounds overlap, sorted and unique (dynamic_aabb_tree.cpp:268-336); fat
// bounds only prune.  An LBVH over tight bounds, rebuilt per step, returns the
// same set.  Build (Karras 2012):
//   1. 24-bit Morton code of each AABB centre in the scene's centre bounds,
//      radix sorted with the index as tie-break;
//   2. one thread per internal node finds its key range and split with
//      binary searches over common-prefix lengths (index bits break ties of
//      equal codes);
//   3. node bounds: an internal node covers a contiguous range of sorted
//      leaves, so its box is a range union built from precomputed boxes of
//      aligned 32 / 1024 / 32768-leaf groups; no atomics or fences (a
//      bottom-up refit would need device-scope fences between XCDs).
// Query: a stackless depth-first walk by ropes.  Node k's children are the
// split's two sides, so the subtree after one whose leaves end at l
// (l < n - 1) in depth-first order is the right child of the node that splits
// l | l + 1: the build records that child per split position, and the walk
// goes there after a subtree it skips or finishes (the last leaf's rope ends
// the walk).  Depth-first order is leaf order, so the walk of a run of
// leaves starts at the rope of its last leaf: it covers exactly the leaves
// after the run (each pair is found once, by its lower sorted position),
// descending into a subtree only when its box meets a box of the run.  Runs
// of four leaves share one walk (k_query).  Children and ropes are encoded
// with bit 31 set for leaves; kEnd ends the walk.
#include <stddef.h>

#include <algorithm>

#include "devmath.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace pg {
namespace {

constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kEnd = 0xFFFFFFFFu;

// Morton bits per axis: 8 (24-bit codes, three 8-bit radix passes).  Equal
// codes are ordered by index, so the tree stays valid with any precision;
// at 1M mixed bodies 10 bits per axis (four passes) give the same query time
// and a 20 us longer sort.
constexpr uint32_t kMortonAxisBits = 8;
static_assert(kMortonAxisBits <= 10, "30-bit Morton codes at most");

__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ __launch_bounds__(kBlock) void k_morton(uint32_t n, const float4* __restrict__ aabb,
                                                   const Counters* counters,
                                                   uint64_t* __restrict__ items)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t code = 0;
    const Rec r = load_rec(aabb, i);
    for (int a = 0; a < 3; ++a) {
        const float lo = key_float(counters->cmin_key[a]);
        const float hi = key_float(counters->cmax_key[a]);
        const float c = r.b[a] * 0.5f + r.b[3 + a] * 0.5f;
        const float ext = hi - lo;
        float u = ext > 0.0f ? (c - lo) / ext : 0.0f;
        u = fminf(fmaxf(u * float(1u << kMortonAxisBits), 0.0f),
                  float((1u << kMortonAxisBits) - 1u));
        code |= spread10(uint32_t(u)) << a;
    }
    items[i] = (uint64_t)code | ((uint64_t)i << 32);
}

// Common-prefix length of sorted keys i and j (-1 outside [0, n)).
__device__ __forceinline__ int delta(const uint64_t* __restrict__ items, int n, int i, int j)
{
    if (j < 0 || j >= n)
        return -1;
    const uint32_t a = uint32_t(items[i]), b = uint32_t(items[j]);
    if (a == b)
        return 32 + __clz(uint32_t(i) ^ uint32_t(j));
    return __clz(a ^ b);
}

__global__ __launch_bounds__(kBlock) void k_build(uint32_t n, const uint64_t* __restrict__ items,
                                                  int32_t* __restrict__ left,
                                                  uint32_t* __restrict__ rope_of_split,
                                                  int32_t* __restrict__ parent,
                                                  uint32_t* __restrict__ first_out,
                                                  uint32_t* __restrict__ last)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ni = int(n);
    if (i >= ni - 1)
        return;
    const int d = (delta(items, ni, i, i + 1) - delta(items, ni, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(items, ni, i, i - d);
    int lmax = 2;
    while (delta(items, ni, i, i + lmax * d) > dmin)
        lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(items, ni, i, i + (l + t) * d) > dmin)
            l += t;
    const int j = i + l * d;
    const int dnode = delta(items, ni, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(items, ni, i, i + (s + t) * d) > dnode)
            s += t;
    } while (t > 1);
    const int split = i + s * d + min(d, 0);
    const int first = min(i, j), lastpos = max(i, j);
    const int32_t lc = (first == split) ? int32_t(uint32_t(split) | kLeaf) : split;
    const int32_t rc = (lastpos == split + 1) ? int32_t(uint32_t(split + 1) | kLeaf) : split + 1;
    left[i] = lc;
    rope_of_split[split] = uint32_t(rc);
    first_out[i] = uint32_t(first);
    last[i] = uint32_t(lastpos);
    // Parents: internal nodes in [0, n-1), leaves in [n-1, 2n-1).
    parent[(lc & int32_t(kLeaf)) ? (ni - 1) + (lc & 0x7FFFFFFF) : lc] = i;
    parent[(rc & int32_t(kLeaf)) ? (ni - 1) + (rc & 0x7FFFFFFF) : rc] = i;
}

// Node bounds without inter-thread synchronisation: an internal node covers
// the sorted leaves [first, last], so its box is the union of that range,
// assembled greedily from the boxes of aligned blocks of 4^k leaves (at most
// three blocks per level on each side, a handful for the common short
// ranges).  min/max are exact, so the order of the unions does not change the
// box.  Level k (1 <= k <= levels) holds count[k] boxes from level_box +
// 2 offset[k].
struct LevelTable {
    uint32_t offset[kLbvhMaxLevels + 1];
    uint32_t count[kLbvhMaxLevels + 1];
    uint32_t levels;
};

__device__ __forceinline__ void empty_box(float u[6])
{
    for (int k = 0; k < 3; ++k) {
        u[k] = __int_as_float(0x7f800000);
        u[3 + k] = -__int_as_float(0x7f800000);
    }
}

__device__ __forceinline__ void unite(float u[6], const float4* __restrict__ r, size_t i)
{
    const Rec b = load_rec(r, i);
    for (int k = 0; k < 3; ++k) {
        u[k] = min_ref(u[k], b.b[k]);
        u[3 + k] = max_ref(u[3 + k], b.b[3 + k]);
    }
}

// Up to five levels above `in` (boxes of level first_level - 1, count_in of
// them): each workgroup reduces 1024 consecutive input boxes to 256, 64, 16,
// 4 and 1 boxes of levels first_level .. first_level + 4.
__global__ __launch_bounds__(256) void k_levels(const float4* __restrict__ in, uint32_t count_in,
                                                float4* __restrict__ base, LevelTable lt,
                                                uint32_t first_level)
{
    __shared__ float box[256][6];
    const uint32_t t = threadIdx.x;
    const uint64_t i0 = (uint64_t(blockIdx.x) * 256 + t) * 4;
    float b[6];
    empty_box(b);
    for (uint32_t j = 0; j < 4; ++j)
        if (i0 + j < count_in)
            unite(b, in, i0 + j);
    uint32_t lev = first_level;
    uint64_t idx = uint64_t(blockIdx.x) * 256 + t;
    if (lev <= lt.levels && idx < lt.count[lev])
        store_rec(base + 2 * size_t(lt.offset[lev]), idx, b, 0);
    for (uint32_t width = 64; width >= 1; width /= 4) {
        for (int k = 0; k < 6; ++k)
            box[t][k] = b[k];
        __syncthreads();
        if (t < width) {
            empty_box(b);
            for (uint32_t j = 0; j < 4; ++j)
                for (int k = 0; k < 3; ++k) {
                    b[k] = min_ref(b[k], box[4 * t + j][k]);
                    b[3 + k] = max_ref(b[3 + k], box[4 * t + j][3 + k]);
                }
        }
        __syncthreads();
        ++lev;
        idx = uint64_t(blockIdx.x) * width + t;
        if (t < width && lev <= lt.levels && idx < lt.count[lev])
            store_rec(base + 2 * size_t(lt.offset[lev]), idx, b, 0);
    }
}

// Node records: the box, the left child (index word) and the rope (world
// word), so one 32-byte read serves a visit.
__global__ __launch_bounds__(kBlock) void k_node_boxes(uint32_t n, const float4* __restrict__ leaves,
                                                       const float4* __restrict__ levels,
                                                       LevelTable lt,
                                                       const uint32_t* __restrict__ first,
                                                       const uint32_t* __restrict__ last,
                                                       const int32_t* __restrict__ left,
                                                       const uint32_t* __restrict__ rope_of_split,
                                                       float4* __restrict__ nodes)
{
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node + 1 >= n)
        return;
    uint64_t i = first[node];
    const uint64_t l = last[node];
    float u[6];
    empty_box(u);
    while (i <= l) {
        // The largest aligned block starting at i that ends within the range.
        uint32_t k = 0;
        while (k < lt.levels && (i & ((uint64_t(4) << (2 * k)) - 1)) == 0 &&
               i + (uint64_t(4) << (2 * k)) - 1 <= l)
            ++k;
        if (k == 0)
            unite(u, leaves, i);
        else
            unite(u, levels + 2 * size_t(lt.offset[k]), i >> (2 * k));
        i += uint64_t(1) << (2 * k);
    }
    const uint32_t end = last[node];
    store_rec(nodes, node, u, uint32_t(left[node]), end + 1 < n ? rope_of_split[end] : kEnd);
}

// Query by packets: a wavefront owns 64 consecutive sorted leaves, one per
// lane, in packets of kPacket lanes.  Pairs inside a packet are tested lane
// against lane.  The leaves after a packet are reached by one walk that the
// packet's lanes share, from the rope of the packet's last leaf, descending
// into a subtree when its box meets the box of any lane of the packet (every
// node above a leaf that meets a lane's box meets that box too); every leaf
// the walk reaches is tested by each lane of the packet.  A packet's walk is
// the union of its lanes' walks: one node read serves the packet's lanes, and
// the wavefront's loop runs as long as its longest packet walk.  (A test
// against the union of the packet's boxes instead of each lane's would let a
// packet that straddles a jump of the Morton curve walk a large part of the
// scene.)
// Four lanes to a packet: per packet about 28 dependent node reads at 1M
// mixed bodies (16 per lane alone, 120 for a whole wavefront), while a
// wavefront issues 16 reads per step instead of 64.
constexpr uint32_t kPacket = 4;
static_assert(kPacket >= 1 && kPacket <= 64 && (kPacket & (kPacket - 1)) == 0,
              "packet: a power of two up to a wavefront");
constexpr uint32_t kQueryBlock = kBlock;

// A wavefront's 16 packet slots take the 32 packets (128 leaves) of its
// range in turn: a slot whose walk has ended takes the range's next packet at
// once, so the wavefront runs about as long as its slots' summed walks rather
// than its longest packet in each batch of 16.  At 1M bodies the ranges make
// 8192 wavefronts, one per wave slot of the GPU (16 packets per range: two
// rounds of wavefronts, 64: half the slots idle).
constexpr uint32_t kSlots = 64 / kPacket;
constexpr uint32_t kRangePackets = 32;

__global__ __launch_bounds__(kQueryBlock) void k_query(uint32_t n, const float4* __restrict__ s,
                                                  const uint32_t* __restrict__ rope_of_split,
                                                  const float4* __restrict__ nodes,
                                                  Counters* counters, uint64_t* __restrict__ out,
                                                  uint64_t cap, uint64_t* __restrict__ visits_part,
                                                  uint64_t* __restrict__ leaves_part)
{
    __shared__ uint64_t queue_lds[kQueryBlock / 64][kWaveQueue];
    WaveQueue queue{queue_lds[threadIdx.x >> 6], 0};
    unsigned long long visits = 0, leaves = 0;
    const unsigned long long max_steps = 2ull * n + 64;
    const uint32_t lane = lane_id();
    const uint32_t rank = lane & (kPacket - 1);  // lane within its packet
    const uint32_t head = lane - rank;           // first lane of the packet slot
    const uint32_t slot = lane / kPacket;
    const uint64_t packet_lanes = (kPacket == 64 ? ~0ull : ((1ull << kPacket) - 1)) << head;
    const uint64_t heads_below = __ballot(rank == 0) & lanes_below(head);
    const uint32_t waves = gridDim.x * (kQueryBlock / 64);
    const uint32_t wave0 =
        xcd_block(blockIdx.x, gridDim.x) * (kQueryBlock / 64) + (threadIdx.x >> 6);
    const uint32_t packets = (n + kPacket - 1) / kPacket;
    for (uint32_t range = wave0 * kRangePackets; range < packets; range += waves * kRangePackets) {
        const uint32_t range_end = min(range + kRangePackets, packets);
        uint32_t next = range + kSlots;  // the range's next unassigned packet (uniform)
        uint32_t pk = range + slot;
        bool live = pk < range_end;  // slot holds a packet (slot-uniform)
        bool fresh = live;           // packet just taken: pairs inside it not yet tested
        uint32_t node = kEnd;
        unsigned long long steps = 0;
        Rec me{{0, 0, 0, 0, 0, 0}, 0, 0};
        uint32_t p = 0;
        bool active = false;
        for (;;) {
            // Slots with a fresh packet: its records, the pairs inside it and
            // the start of its walk (the rope of its last leaf).
            if (__ballot(fresh)) {
                if (fresh) {
                    p = pk * kPacket + rank;
                    active = p < n;
                    me = load_rec(s, active ? p : n - 1);
                    const uint32_t last = min(pk * kPacket + kPacket, n) - 1;
                    node = last + 1 < n ? rope_of_split[last] : kEnd;
                    steps = 0;
                }
                for (uint32_t j = 1; j < kPacket; ++j) {
                    const int src = int(lane + j < 64 ? lane + j : lane);
                    const float ob[6] = {__shfl(me.b[0], src), __shfl(me.b[1], src),
                                         __shfl(me.b[2], src), __shfl(me.b[3], src),
                                         __shfl(me.b[4], src), __shfl(me.b[5], src)};
                    const uint32_t oi = __shfl(me.index, src);
                    const bool valid = fresh && rank + j < kPacket && p + j < n;
                    leaves += valid ? 1u : 0u;
                    wq_push(queue, valid && rec_overlap(me.b, ob), pair_value(me.index, oi), out,
                            &counters->pair_count, cap);
                }
                fresh = false;
            }
            // One step of every live walk.  A walk enters each node at most
            // once; the cap only guards against a malformed tree.
            if (node != kEnd && steps > max_steps) {
                if (rank == 0)
                    atomicAdd(&counters->tree_overflow, 1u);
                node = kEnd;
            }
            bool hit = false;
            uint32_t other = 0;
            if (node != kEnd) {
                ++steps;
                if (node & kLeaf) {
                    const uint32_t q = node & ~kLeaf;
                    const Rec o = load_rec(s, q);
                    node = q + 1 < n ? rope_of_split[q] : kEnd;
                    leaves += active ? 1u : 0u;
                    hit = active && rec_overlap(me.b, o.b);
                    other = o.index;
                } else {
                    const Rec nd = load_rec(nodes, node);
                    visits += rank == 0 ? 1u : 0u;
                    // Every lane of a packet takes this branch together.
                    const bool meets = (__ballot(active && rec_overlap(me.b, nd.b)) &
                                        packet_lanes) != 0;
                    node = meets ? nd.index : nd.world;
                }
            }
            wq_push(queue, hit, pair_value(me.index, other), out, &counters->pair_count, cap);
            // Slots whose walk ended take the range's next packets, in slot
            // order; with none left they fall idle.
            const bool done = live && node == kEnd;
            const uint64_t done_heads = __ballot(done && rank == 0);
            if (done_heads) {
                const uint32_t k = uint32_t(__popcll(done_heads & heads_below));
                if (done) {
                    pk = next + k;
                    live = pk < range_end;
                    fresh = live;
                }
                next += uint32_t(__popcll(done_heads));
            }
            if (__ballot(live) == 0)
                break;
        }
    }
    wq_flush_block(queue, out, &counters->pair_count, cap);
    block_partial(visits, visits_part);
    block_partial(leaves, leaves_part);
}

}  // namespace

void launch_lbvh(physgpu_context* ctx)
{
    const uint32_t n = ctx->n;
    Timer t;
    time_begin(ctx, "lbvh_morton", t);
    hipLaunchKernelGGL(k_morton, dim3(div_up(n, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       ctx->aabb, ctx->counters, ctx->items_a);
    time_end(ctx, t);
    uint64_t* sorted = nullptr;
    radix_sort(ctx, ctx->items_a, ctx->items_b, n, nullptr, 0, 3 * kMortonAxisBits, 0, 0, &sorted,
               ctx->aabb,
               ctx->sorted_aabb);
    ctx->lbvh_sorted = sorted;
    if (n < 2)
        return;
    time_begin(ctx, "lbvh_build", t);
    hipLaunchKernelGGL(k_build, dim3(div_up(n - 1, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       sorted, ctx->node_left, reinterpret_cast<uint32_t*>(ctx->node_right),
                       ctx->node_parent, ctx->node_first, ctx->node_last);
    time_end(ctx, t);
    launch_lbvh_boxes_and_query(ctx);
}

// Refit: the leaf order and the nodes of the last build, with the records'
// current bounds gathered in that order and every node box recomputed.  The
// query stays exact (boxes bound their subtrees); only pruning can loosen.
void launch_lbvh_refit(physgpu_context* ctx)
{
    launch_gather_sorted(ctx, ctx->lbvh_sorted);
    if (ctx->n < 2)
        return;
    launch_lbvh_boxes_and_query(ctx);
}

void launch_lbvh_boxes_and_query(physgpu_context* ctx)
{
    const uint32_t n = ctx->n;
    Timer t;
    // Level boxes of aligned 4^k-leaf blocks, five levels per launch.
    LevelTable lt{};
    uint32_t total = 0, count = n;
    while (count > 1 && lt.levels < kLbvhMaxLevels) {
        count = div_up(count, 4u);
        ++lt.levels;
        lt.offset[lt.levels] = total;
        lt.count[lt.levels] = count;
        total += count;
    }
    time_begin(ctx, "lbvh_node_boxes", t);
    for (uint32_t from = 1; from <= lt.levels; from += 5) {
        const float4* in = from == 1 ? ctx->sorted_aabb
                                     : ctx->level_box + 2 * size_t(lt.offset[from - 1]);
        const uint32_t count_in = from == 1 ? n : lt.count[from - 1];
        hipLaunchKernelGGL(k_levels, dim3(div_up(count_in, 1024u)), dim3(256), 0, ctx->stream, in,
                           count_in, ctx->level_box, lt, from);
    }
    hipLaunchKernelGGL(k_node_boxes, dim3(div_up(n - 1, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       ctx->sorted_aabb, ctx->level_box, lt, ctx->node_first, ctx->node_last,
                       ctx->node_left, reinterpret_cast<const uint32_t*>(ctx->node_right),
                       ctx->node_box);
    time_end(ctx, t);
    time_begin(ctx, "lbvh_query", t);
    const uint32_t qgrid = std::min<uint32_t>(
        div_up(div_up(div_up(n, kPacket), kRangePackets), kQueryBlock / 64), kMaxPartialBlocks);
    hipLaunchKernelGGL(k_query, dim3(qgrid), dim3(kQueryBlock), 0, ctx->stream, n,
                       ctx->sorted_aabb, reinterpret_cast<const uint32_t*>(ctx->node_right),
                       ctx->node_box, ctx->counters, ctx->pairs_raw, ctx->cap_p,
                       ctx->partials + size_t(kSlotTreeVisits) * kMaxPartialBlocks,
                       ctx->partials + size_t(kSlotTreeLeaves) * kMaxPartialBlocks);
    time_end(ctx, t);
    FoldList fold{};
    fold.e[0] = {kSlotTreeVisits, qgrid, kFoldSum64, uint32_t(offsetof(Counters, tree_visits))};
    fold.e[1] = {kSlotTreeLeaves, qgrid, kFoldSum64,
                 uint32_t(offsetof(Counters, tree_leaf_checks))};
    fold.count = 2;
    launch_fold(ctx, fold);
}

}  // namespace pg
