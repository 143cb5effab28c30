// Linear BVH standing in for DynamicAabbTree (src/collision/dynamic_aabb_tree.cpp).
//
// The reference tree's query returns every pair of proxies whose *tight*
// bounds overlap, sorted and unique (dynamic_aabb_tree.cpp:268-336); fat
// bounds only prune.  An LBVH over tight bounds, rebuilt per step, returns the
// same set.  Build (Karras 2012):
//   1. 30-bit Morton code of each AABB centre in the scene's centre bounds,
//      radix sorted with the index as tie-break;
//   2. one thread per internal node finds its key range and split with
//      binary searches over common-prefix lengths (index bits break ties of
//      equal codes);
//   3. node bounds: an internal node covers a contiguous range of sorted
//      leaves, so its box is a range union built from precomputed boxes of
//      aligned 32 / 1024 / 32768-leaf groups; no atomics or fences (a
//      bottom-up refit would need device-scope fences between XCDs).
// Query: one thread per leaf, stack traversal; a subtree is skipped when all
// of its leaves sit at or before the querying leaf in sorted order, so each
// pair is found once.  Children are encoded with bit 31 set for leaves; a
// stack entry with bit 31 is an overlapping leaf, emitted when popped (it is
// never used to index the internal-node arrays).
#include <stddef.h>

#include <algorithm>

#include "devmath.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace pg {
namespace {

constexpr uint32_t kLeaf = 0x80000000u;
constexpr int kStack = 96;  // > 62-bit key depth + pending siblings

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
        u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
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
                                                  int32_t* __restrict__ right,
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
    right[i] = rc;
    first_out[i] = uint32_t(first);
    last[i] = uint32_t(lastpos);
    // Parents: internal nodes in [0, n-1), leaves in [n-1, 2n-1).
    parent[(lc & int32_t(kLeaf)) ? (ni - 1) + (lc & 0x7FFFFFFF) : lc] = i;
    parent[(rc & int32_t(kLeaf)) ? (ni - 1) + (rc & 0x7FFFFFFF) : rc] = i;
}

// Node bounds without inter-thread synchronisation: an internal node covers
// the sorted leaves [first, last], so its box is the union of that range,
// assembled from aligned blocks of 32^3, 32^2, 32 leaves and single leaves.
// (min/max are exact, so the order of the unions does not change the box.)
constexpr uint32_t kSpan = 32;

__device__ __forceinline__ void unite(float u[6], const float4* __restrict__ r, size_t i)
{
    const Rec b = load_rec(r, i);
    for (int k = 0; k < 3; ++k) {
        u[k] = min_ref(u[k], b.b[k]);
        u[3 + k] = max_ref(u[3 + k], b.b[3 + k]);
    }
}

// Boxes of consecutive groups of 32 records (one per half wave).
__global__ __launch_bounds__(kBlock) void k_group_boxes(uint32_t count, const float4* __restrict__ in,
                                                        float4* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = i < count;
    float b[6];
    if (ok) {
        const Rec r = load_rec(in, i);
        for (int k = 0; k < 6; ++k)
            b[k] = r.b[k];
    } else {
        for (int k = 0; k < 3; ++k) {
            b[k] = __int_as_float(0x7f800000);
            b[3 + k] = -__int_as_float(0x7f800000);
        }
    }
    for (uint32_t off = 1; off < kSpan; off <<= 1)
        for (int k = 0; k < 3; ++k) {
            b[k] = fminf(b[k], __shfl_xor(b[k], off));
            b[3 + k] = fmaxf(b[3 + k], __shfl_xor(b[3 + k], off));
        }
    if ((i % kSpan) == 0 && ok)
        store_rec(out, i / kSpan, b, 0);
}

__global__ __launch_bounds__(kBlock) void k_node_boxes(uint32_t n, const float4* __restrict__ leaves,
                                                       const float4* __restrict__ l1,
                                                       const float4* __restrict__ l2,
                                                       const float4* __restrict__ l3,
                                                       const uint32_t* __restrict__ first,
                                                       const uint32_t* __restrict__ last,
                                                       float4* __restrict__ nodes)
{
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node + 1 >= n)
        return;
    uint32_t i = first[node];
    const uint32_t l = last[node];
    float u[6];
    for (int k = 0; k < 3; ++k) {
        u[k] = __int_as_float(0x7f800000);
        u[3 + k] = -__int_as_float(0x7f800000);
    }
    constexpr uint32_t s1 = kSpan, s2 = kSpan * kSpan, s3 = kSpan * kSpan * kSpan;
    while (i <= l) {
        if (i % s3 == 0 && uint64_t(i) + s3 - 1 <= l) {
            unite(u, l3, i / s3);
            i += s3;
        } else if (i % s2 == 0 && uint64_t(i) + s2 - 1 <= l) {
            unite(u, l2, i / s2);
            i += s2;
        } else if (i % s1 == 0 && uint64_t(i) + s1 - 1 <= l) {
            unite(u, l1, i / s1);
            i += s1;
        } else {
            unite(u, leaves, i);
            i += 1;
        }
    }
    store_rec(nodes, node, u, 0);
}

__global__ __launch_bounds__(kBlock) void k_query(uint32_t n, const float4* __restrict__ s,
                                                  const int32_t* __restrict__ left,
                                                  const int32_t* __restrict__ right,
                                                  const uint32_t* __restrict__ last,
                                                  const float4* __restrict__ nodes,
                                                  Counters* counters, uint64_t* __restrict__ out,
                                                  uint64_t cap, uint64_t* __restrict__ visits_part,
                                                  uint64_t* __restrict__ leaves_part)
{
    __shared__ uint64_t queue_lds[kBlock / 64][kWaveQueue];
    WaveQueue queue{queue_lds[threadIdx.x >> 6], 0};
    unsigned long long visits = 0, leaves = 0;
    const unsigned long long max_visits = 2ull * n + 64;
    // Grid-stride over leaves; every lane stays in the loops so the
    // wave-wide appends run converged.
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;
    const uint32_t block = xcd_block(blockIdx.x, gridDim.x);
    for (uint32_t round = 0; round < rounds; ++round) {
        const uint32_t p = round * stride + block * blockDim.x + threadIdx.x;
        unsigned long long my_visits = 0;
        int32_t stack[kStack];
        int sp = 0;
        Rec me{{0, 0, 0, 0, 0, 0}, 0, 0};
        const bool active = p < n && n >= 2;
        if (active) {
            me = load_rec(s, p);
            stack[sp++] = 0;  // root
        }
        // Converged loop: every lane pops until all stacks are empty.  A lane
        // never visits more nodes than the tree has; the cap only guards
        // against a malformed tree.
        for (;;) {
            if (my_visits > max_visits) {
                atomicAdd(&counters->tree_overflow, 1u);
                sp = 0;
            }
            const bool more = sp > 0;
            if (__ballot(more) == 0)
                break;
            bool hit = false;
            uint32_t b = 0;
            if (more) {
                const int32_t entry = stack[--sp];
                if (entry & int32_t(kLeaf)) {
                    // A leaf pushed below already passed the order and overlap tests.
                    hit = true;
                    b = load_rec(s, uint32_t(entry) & 0x7FFFFFFFu).index;
                } else {
                    ++my_visits;
                    const int32_t kids[2] = {left[entry], right[entry]};
                    for (int c = 0; c < 2; ++c) {
                        const int32_t child = kids[c];
                        bool take = false;
                        if (child & int32_t(kLeaf)) {
                            const uint32_t q = uint32_t(child) & 0x7FFFFFFFu;
                            if (q <= p)
                                continue;
                            ++leaves;
                            take = rec_overlap(me.b, load_rec(s, q).b);
                        } else {
                            if (last[child] <= p)
                                continue;
                            take = rec_overlap(me.b, load_rec(nodes, uint32_t(child)).b);
                        }
                        if (take) {
                            if (sp < kStack)
                                stack[sp++] = child;
                            else
                                atomicAdd(&counters->tree_overflow, 1u);
                        }
                    }
                }
            }
            wq_push(queue, hit, pair_value(me.index, b), out, &counters->pair_count, cap);
        }
        visits += my_visits;
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
    radix_sort(ctx, ctx->items_a, ctx->items_b, n, nullptr, 0, 30, 0, 0, &sorted, ctx->aabb,
               ctx->sorted_aabb);
    ctx->lbvh_sorted = sorted;
    if (n < 2)
        return;
    time_begin(ctx, "lbvh_build", t);
    hipLaunchKernelGGL(k_build, dim3(div_up(n - 1, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       sorted, ctx->node_left, ctx->node_right, ctx->node_parent, ctx->node_first,
                       ctx->node_last);
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
    // Group boxes: 32, 1024 and 32768 leaves.
    const uint32_t c1 = div_up(n, kSpan), c2 = div_up(c1, kSpan), c3 = div_up(c2, kSpan);
    float4* l1 = ctx->level_box;
    float4* l2 = l1 + 2 * size_t(c1);
    float4* l3 = l2 + 2 * size_t(c2);
    time_begin(ctx, "lbvh_node_boxes", t);
    hipLaunchKernelGGL(k_group_boxes, dim3(div_up(n, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       ctx->sorted_aabb, l1);
    hipLaunchKernelGGL(k_group_boxes, dim3(div_up(c1, kBlock)), dim3(kBlock), 0, ctx->stream, c1,
                       l1, l2);
    hipLaunchKernelGGL(k_group_boxes, dim3(div_up(c2, kBlock)), dim3(kBlock), 0, ctx->stream, c2,
                       l2, l3);
    hipLaunchKernelGGL(k_node_boxes, dim3(div_up(n - 1, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       ctx->sorted_aabb, l1, l2, l3, ctx->node_first, ctx->node_last,
                       ctx->node_box);
    time_end(ctx, t);
    (void)c3;
    time_begin(ctx, "lbvh_query", t);
    const uint32_t qgrid = std::min<uint32_t>(div_up(n, kBlock), 4096);
    hipLaunchKernelGGL(k_query, dim3(qgrid), dim3(kBlock), 0, ctx->stream, n,
                       ctx->sorted_aabb, ctx->node_left, ctx->node_right, ctx->node_last,
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
