// Batches of independent worlds in one context: the resident colliders are
// the concatenation of the worlds, world w owning [offsets[w], offsets[w+1]).
// Each world keeps the per-world semantics of PhysicsWorld::step
// (src/world/physics_world.cpp:518-782): pairs only form inside a world, and
// because worlds occupy increasing index ranges the canonical order of the
// whole pair list is the concatenation of the per-world canonical lists.
// These kernels tag colliders with their world and cut the step's pair and
// manifold lists into per-world ranges.
#include "internal.hpp"

namespace pg {
namespace {

// World of every collider: the last w with offsets[w] <= i (binary search).
__global__ __launch_bounds__(kBlock) void k_world_of(uint32_t n, uint32_t worlds,
                                                     const uint32_t* __restrict__ offsets,
                                                     uint32_t* __restrict__ world_of)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t lo = 0, hi = worlds;  // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    world_of[i] = lo;
}

// First position in sorted[0, count) whose value is >= key.
template <typename T, typename Key>
__device__ __forceinline__ uint32_t lower_bound(const T* __restrict__ sorted, uint32_t count,
                                                uint64_t key, Key key_of)
{
    uint32_t lo = 0, hi = count;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key_of(sorted[mid]) < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Per-world ranges of the canonical pair list and of the compacted manifold
// list (entry `worlds` = the totals).
__global__ __launch_bounds__(kBlock) void k_world_ranges(
    uint32_t worlds, const uint32_t* __restrict__ offsets, const uint64_t* __restrict__ pairs,
    const physgpu_manifold* __restrict__ manifolds, const Counters* __restrict__ c,
    uint64_t* __restrict__ pair_off, uint64_t* __restrict__ manifold_off)
{
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w > worlds)
        return;
    const uint32_t held = c->sort_count, made = c->manifold_count;
    if (w == worlds) {
        pair_off[w] = held;
        manifold_off[w] = made;
        return;
    }
    const uint64_t first = offsets[w];
    pair_off[w] = lower_bound(pairs, held, first << 32, [](uint64_t v) { return v; });
    manifold_off[w] = lower_bound(manifolds, made, first,
                                  [](const physgpu_manifold& m) { return uint64_t(m.first); });
}

}  // namespace

void launch_world_of(physgpu_context* ctx)
{
    if (ctx->n == 0 || ctx->n_worlds == 0)
        return;
    hipLaunchKernelGGL(k_world_of, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->n, ctx->n_worlds, ctx->world_offsets, ctx->world_of);
}

void launch_world_ranges(physgpu_context* ctx)
{
    if (ctx->n_worlds == 0)
        return;
    hipLaunchKernelGGL(k_world_ranges, dim3(div_up(ctx->n_worlds + 1, kBlock)), dim3(kBlock), 0,
                       ctx->stream, ctx->n_worlds, ctx->world_offsets, ctx->pairs, ctx->manifolds,
                       ctx->counters, ctx->world_pair_off, ctx->world_manifold_off);
}

}  // namespace pg
