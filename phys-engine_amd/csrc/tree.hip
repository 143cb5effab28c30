// Persistent proxies (DynamicAabbTree, include/phys/collision/dynamic_aabb_tree.hpp:25-31,
// src/collision/dynamic_aabb_tree.cpp:44-57, 108-139, 268-336) over the LBVH.
//
// Device planes indexed by proxy id hold each proxy's tight and fat bounds in
// the AABB record layout (2 float4).  The fat-bounds rule is the reference's:
// created and reinserted proxies get expanded(bounds) (padding max(0.01, 10% of
// the side) per axis, in double, clamped to the float range); an update keeps
// the fat bounds while they contain the new tight bounds and expanded(new, 4)
// contains them, else reinserts.  The pair set of the reference's query is
// every pair of live proxies whose tight bounds overlap (fat bounds contain
// the tight ones, so fat pruning never drops such a pair); it is found by the
// LBVH over the tight bounds, laid out by the rank of each proxy's user index
// among the live proxies, so that canonical pairs in rank space map to user
// indices in canonical order.
#include <cfloat>

#include "internal.hpp"
#include "wave.hpp"

namespace pg {
namespace {

// expanded(bounds, factor), dynamic_aabb_tree.cpp:44-57.
__device__ __forceinline__ void expanded(const float b[6], double factor, float out[6])
{
    const double limit = double(FLT_MAX);
    for (int a = 0; a < 3; ++a) {
        const double lo = double(b[a]), hi = double(b[3 + a]);
        double pad = (hi - lo) * 0.1;
        pad = 0.01 < pad ? pad : 0.01;
        pad = pad * factor;
        const double l = lo - pad, h = hi + pad;
        out[a] = float(-limit < l ? l : -limit);
        out[3 + a] = float(h < limit ? h : limit);
    }
}

__device__ __forceinline__ bool contains(const float outer[6], const float inner[6])
{
    return outer[0] <= inner[0] && outer[3] >= inner[3] && outer[1] <= inner[1] &&
           outer[4] >= inner[4] && outer[2] <= inner[2] && outer[5] >= inner[5];
}

__device__ __forceinline__ void load_bounds(const physgpu_aabb& a, float b[6])
{
    b[0] = a.min[0];
    b[1] = a.min[1];
    b[2] = a.min[2];
    b[3] = a.max[0];
    b[4] = a.max[1];
    b[5] = a.max[2];
}

// createProxy (dynamic_aabb_tree.cpp:108-118): tight = bounds, fat = expanded.
__global__ __launch_bounds__(kBlock) void k_tree_create(uint32_t count, const uint32_t* __restrict__ ids,
                                                        const physgpu_aabb* __restrict__ bounds,
                                                        float4* __restrict__ tight,
                                                        float4* __restrict__ fat)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const uint32_t p = ids[i];
    float b[6], f[6];
    load_bounds(bounds[i], b);
    expanded(b, 1.0, f);
    store_rec(tight, p, b, p);
    store_rec(fat, p, f, p);
}

// updateProxy (dynamic_aabb_tree.cpp:128-139); reinserted[i] = 1 when the
// proxy would be reinserted.
__global__ __launch_bounds__(kBlock) void k_tree_update(uint32_t count, const uint32_t* __restrict__ ids,
                                                        const physgpu_aabb* __restrict__ bounds,
                                                        float4* __restrict__ tight,
                                                        float4* __restrict__ fat,
                                                        uint8_t* __restrict__ reinserted)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const uint32_t p = ids[i];
    float b[6], wide[6];
    load_bounds(bounds[i], b);
    store_rec(tight, p, b, p);
    const Rec f = load_rec(fat, p);
    expanded(b, 4.0, wide);
    const bool keep = contains(f.b, b) && contains(wide, f.b);
    if (!keep) {
        float nf[6];
        expanded(b, 1.0, nf);
        store_rec(fat, p, nf, p);
    }
    reinserted[i] = keep ? 0 : 1;
}

// AABB records of the live proxies at their ranks (the broad-phase input).
__global__ __launch_bounds__(kBlock) void k_tree_records(uint32_t slots, const uint32_t* __restrict__ rank,
                                                         const float4* __restrict__ tight,
                                                         float4* __restrict__ records)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= slots)
        return;
    const uint32_t r = rank[p];
    if (r == 0xFFFFFFFFu)
        return;
    const Rec t = load_rec(tight, p);
    store_rec(records, r, t.b, r);
}

// Canonical rank pairs to user-index pairs; ranks grow with user indices, so
// the order stays canonical.
__global__ __launch_bounds__(kBlock) void k_tree_pairs_out(const uint64_t* __restrict__ pairs,
                                                           const Counters* __restrict__ c,
                                                           const uint64_t* __restrict__ rank_user,
                                                           physgpu_proxy_pair* __restrict__ out)
{
    const uint32_t count = c->sort_count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += gridDim.x * blockDim.x) {
        const uint64_t v = pairs[i];
        out[i] = {rank_user[uint32_t(v >> 32)], rank_user[uint32_t(v)]};
    }
}

}  // namespace

void launch_tree_create(physgpu_context* ctx, uint32_t count, const uint32_t* ids,
                        const physgpu_aabb* bounds)
{
    if (count)
        hipLaunchKernelGGL(k_tree_create, dim3(div_up(count, kBlock)), dim3(kBlock), 0,
                           ctx->stream, count, ids, bounds, ctx->tree_tight, ctx->tree_fat);
}

void launch_tree_update(physgpu_context* ctx, uint32_t count, const uint32_t* ids,
                        const physgpu_aabb* bounds, uint8_t* reinserted)
{
    if (count)
        hipLaunchKernelGGL(k_tree_update, dim3(div_up(count, kBlock)), dim3(kBlock), 0,
                           ctx->stream, count, ids, bounds, ctx->tree_tight, ctx->tree_fat,
                           reinserted);
}

void launch_tree_records(physgpu_context* ctx, uint32_t slots)
{
    if (slots)
        hipLaunchKernelGGL(k_tree_records, dim3(div_up(slots, kBlock)), dim3(kBlock), 0,
                           ctx->stream, slots, ctx->tree_rank, ctx->tree_tight, ctx->aabb);
}

void launch_tree_pairs_out(physgpu_context* ctx, physgpu_proxy_pair* out)
{
    hipLaunchKernelGGL(k_tree_pairs_out, dim3(stride_grid(ctx->cap_p)), dim3(kBlock), 0,
                       ctx->stream, ctx->pairs, ctx->counters, ctx->tree_rank_user, out);
}

}  // namespace pg
