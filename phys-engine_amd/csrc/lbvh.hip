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
//   3. bottom-up refit: each leaf climbs, the second arrival at a node merges
//      its children (device-scope fences make the first child's box visible).
// Query: one thread per leaf, stack traversal; a subtree is skipped when all
// of its leaves sit at or before the querying leaf in sorted order, so each
// pair is found once.  Children are encoded with bit 31 set for leaves; a
// stack entry with bit 31 is an overlapping leaf, emitted when popped (it is
// never used to index the internal-node arrays).
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

__global__ __launch_bounds__(kBlock) void k_morton(uint32_t n, const float* __restrict__ aabb,
                                                   const Counters* counters,
                                                   uint64_t* __restrict__ items)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t code = 0;
    for (int a = 0; a < 3; ++a) {
        const float lo = key_float(counters->cmin_key[a]);
        const float hi = key_float(counters->cmax_key[a]);
        const float c = aabb[a * n + i] * 0.5f + aabb[(3 + a) * n + i] * 0.5f;
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
                                                  uint32_t* __restrict__ last,
                                                  uint32_t* __restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ni = int(n);
    if (i >= ni - 1)
        return;
    flag[i] = 0;
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
    last[i] = uint32_t(lastpos);
    // Parents: internal nodes in [0, n-1), leaves in [n-1, 2n-1).
    parent[(lc & int32_t(kLeaf)) ? (ni - 1) + (lc & 0x7FFFFFFF) : lc] = i;
    parent[(rc & int32_t(kLeaf)) ? (ni - 1) + (rc & 0x7FFFFFFF) : rc] = i;
}

__device__ __forceinline__ void load_box(const float* __restrict__ s, size_t n,
                                         const float* __restrict__ nodes, size_t m, int32_t child,
                                         float b[6])
{
    if (child & int32_t(kLeaf)) {
        const size_t p = uint32_t(child) & 0x7FFFFFFFu;
        for (int a = 0; a < 6; ++a)
            b[a] = s[a * n + p];
    } else {
        for (int a = 0; a < 6; ++a)
            b[a] = nodes[a * m + child];
    }
}

__global__ __launch_bounds__(kBlock) void k_refit(uint32_t n, const float* __restrict__ s,
                                                  const int32_t* __restrict__ left,
                                                  const int32_t* __restrict__ right,
                                                  const int32_t* __restrict__ parent,
                                                  uint32_t* __restrict__ flag,
                                                  float* __restrict__ nodes, Counters* counters)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || n < 2)
        return;
    const size_t m = n - 1;
    int32_t node = parent[(n - 1) + p];
    uint32_t depth = 1;
    while (node >= 0) {
        if (depth > 2 * kStack) {  // deeper than any 62-bit-key tree: malformed
            atomicAdd(&counters->tree_overflow, 1u);
            return;
        }
        __threadfence();
        if (atomicAdd(&flag[node], 1u) == 0)
            return;  // the sibling's thread merges this node
        __threadfence();
        float a[6], b[6];
        load_box(s, n, nodes, m, left[node], a);
        load_box(s, n, nodes, m, right[node], b);
        for (int k = 0; k < 3; ++k) {
            nodes[k * m + node] = min_ref(a[k], b[k]);
            nodes[(3 + k) * m + node] = max_ref(a[3 + k], b[3 + k]);
        }
        ++depth;
        node = node == 0 ? -1 : parent[node];
    }
    atomicMax(&counters->tree_height, depth);
}

__device__ __forceinline__ bool overlap6(const float a[6], const float b[6])
{
    return a[0] <= b[3] && a[3] >= b[0] && a[1] <= b[4] && a[4] >= b[1] && a[2] <= b[5] &&
           a[5] >= b[2];
}

__global__ __launch_bounds__(kBlock) void k_query(uint32_t n, const float* __restrict__ s,
                                                  const uint32_t* __restrict__ sidx,
                                                  const int32_t* __restrict__ left,
                                                  const int32_t* __restrict__ right,
                                                  const uint32_t* __restrict__ last,
                                                  const float* __restrict__ nodes,
                                                  Counters* counters, uint64_t* __restrict__ out,
                                                  uint64_t cap)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t m = n - 1;
    unsigned long long visits = 0, leaves = 0;
    int32_t stack[kStack];
    int sp = 0;
    float box[6] = {0, 0, 0, 0, 0, 0};
    uint32_t a = 0;
    const bool active = p < n && n >= 2;
    if (active) {
        for (int k = 0; k < 6; ++k)
            box[k] = s[k * (size_t)n + p];
        a = sidx[p];
        stack[sp++] = 0;  // root
    }
    // Converged loop: every lane pops until all stacks are empty.  A lane
    // never visits more nodes than the tree has; the cap only guards against
    // a malformed tree.
    const unsigned long long max_visits = 2ull * n + 64;
    for (;;) {
        if (visits > max_visits) {
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
                b = sidx[uint32_t(entry) & 0x7FFFFFFFu];
            } else {
                ++visits;
                const int32_t kids[2] = {left[entry], right[entry]};
                for (int c = 0; c < 2; ++c) {
                    const int32_t child = kids[c];
                    float cb[6];
                    bool take = false;
                    if (child & int32_t(kLeaf)) {
                        const uint32_t q = uint32_t(child) & 0x7FFFFFFFu;
                        if (q <= p)
                            continue;
                        for (int k = 0; k < 6; ++k)
                            cb[k] = s[k * (size_t)n + q];
                        ++leaves;
                        take = overlap6(box, cb);
                    } else {
                        if (last[child] <= p)
                            continue;
                        for (int k = 0; k < 6; ++k)
                            cb[k] = nodes[k * m + child];
                        take = overlap6(box, cb);
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
        wave_append(hit, pair_value(a, b), out, &counters->pair_count, cap);
    }
    visits = wave_sum(visits);
    leaves = wave_sum(leaves);
    if (lane_id() == 0) {
        atomicAdd(&counters->tree_visits, visits);
        atomicAdd(&counters->tree_leaf_checks, leaves);
    }
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
    radix_sort(ctx, ctx->items_a, ctx->items_b, n, nullptr, 0, 30, 0, 0, &sorted);
    launch_gather_sorted(ctx, sorted);
    if (n < 2)
        return;
    time_begin(ctx, "lbvh_build", t);
    hipLaunchKernelGGL(k_build, dim3(div_up(n - 1, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       sorted, ctx->node_left, ctx->node_right, ctx->node_parent, ctx->node_last,
                       ctx->node_flag);
    time_end(ctx, t);
    time_begin(ctx, "lbvh_refit", t);
    hipLaunchKernelGGL(k_refit, dim3(div_up(n, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       ctx->sorted_aabb, ctx->node_left, ctx->node_right, ctx->node_parent,
                       ctx->node_flag, ctx->node_box, ctx->counters);
    time_end(ctx, t);
    time_begin(ctx, "lbvh_query", t);
    hipLaunchKernelGGL(k_query, dim3(div_up(n, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       ctx->sorted_aabb, ctx->sorted_index, ctx->node_left, ctx->node_right,
                       ctx->node_last, ctx->node_box, ctx->counters, ctx->pairs_raw, ctx->cap_p);
    time_end(ctx, t);
}

}  // namespace pg
