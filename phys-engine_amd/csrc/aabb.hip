// AABB refresh and per-step reductions.
//
// k_compute_aabbs follows Aabb::fromCollider (src/collision/aabb.cpp:8-32):
//   sphere  pos -/+ (r, r, r)
//   box     pos -/+ |R| h with R = toMat3(q), row sums in the reference order
//   hull    component-wise min/max of pos + rotate(q, p_k) over the points
//           (no hull in the reference; this is the builder's definition)
// One thread per collider; the seven pose planes, type and shape planes are
// read coalesced and one 32-byte AABB record (wave.hpp Rec) written per body.
//
// k_reduce reads the AABB records once and produces what the broad phases
// need: largest x side (SAP row window bound), largest side and mean side
// (grid cell), and centre bounds (LBVH Morton quantisation).  Wave-level
// reductions, one atomic per wave.
#include <stddef.h>

#include "devmath.hpp"
#include "gridkey.hpp"
#include "internal.hpp"
#include "wave.hpp"

#include <algorithm>

namespace pg {
namespace {

// Step counters to their initial values (the calling workgroup's threads).
__device__ __forceinline__ void reset_counters(Counters* c)
{
    uint32_t* w = reinterpret_cast<uint32_t*>(c);
    for (uint32_t i = threadIdx.x; i < sizeof(Counters) / 4; i += blockDim.x)
        w[i] = 0;
    __syncthreads();
    if (threadIdx.x < 3) {
        c->cmin_key[threadIdx.x] = 0xFFFFFFFFu;
        c->cmax_key[threadIdx.x] = 0u;
    }
}

__global__ void k_step_begin(Counters* c)
{
    reset_counters(c);
}

// AABB of sphere or box i (fromCollider, aabb.cpp); false for a hull.
__device__ __forceinline__ bool body_aabb(uint32_t i, const float4* __restrict__ pose,
                                          const float4* __restrict__ shape, float b[6])
{
    const float4 s0 = shape[2 * size_t(i)];
    const uint32_t ty = __float_as_uint(s0.w);
    if (ty != PHYSGPU_SPHERE && ty != PHYSGPU_BOX)
        return false;
    const pg::pose t = load_pose4(pose, i);
    f3 mn, mx;
    if (ty == PHYSGPU_SPHERE) {
        const float r = s0.x;
        f3 rr{r, r, r};
        mn = sub(t.p, rr);
        mx = add(t.p, rr);
    } else {
        const f3 h{s0.x, s0.y, s0.z};
        const m3 m = to_mat(t.q);
        f3 e{fabsf(m.m00) * h.x + fabsf(m.m01) * h.y + fabsf(m.m02) * h.z,
             fabsf(m.m10) * h.x + fabsf(m.m11) * h.y + fabsf(m.m12) * h.z,
             fabsf(m.m20) * h.x + fabsf(m.m21) * h.y + fabsf(m.m22) * h.z};
        mn = sub(t.p, e);
        mx = add(t.p, e);
    }
    b[0] = mn.x;
    b[1] = mn.y;
    b[2] = mn.z;
    b[3] = mx.x;
    b[4] = mx.y;
    b[5] = mx.z;
    return true;
}

__global__ __launch_bounds__(kBlock) void k_compute_aabbs(
    uint32_t n, const float4* __restrict__ pose, const float4* __restrict__ shape,
    const uint32_t* __restrict__ world_of, float4* __restrict__ aabb)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float b[6];
    if (i < n && body_aabb(i, pose, shape, b))  // hulls: k_hull_aabbs
        store_rec(aabb, i, b, i, world_of ? world_of[i] : 0u);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t count, uint32_t value)
{
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x)
        p[i] = value;
}

// Hull bodies (listed once per upload by k_hull_list).
__global__ __launch_bounds__(kBlock) void k_hull_list(uint32_t n, const uint8_t* __restrict__ type,
                                                      uint32_t* __restrict__ list,
                                                      uint32_t* __restrict__ count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool hull = i < n && type[i] == PHYSGPU_HULL;
    const uint64_t mask = __ballot(hull);
    if (!mask)
        return;
    const uint32_t lane = lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    if (hull)
        list[base + (uint32_t)__popcll(mask & lanes_below(lane))] = i;
}

// First extreme (lowest index among equal values, as the sequential min_ref /
// max_ref scan keeps the earlier point) of one component over a 16-lane row
// whose lane holds the first extreme of its points rl, rl + 16, ...: false when
// a value is NaN.
template <bool Max>
__device__ __forceinline__ bool row_first_extreme(float& v, uint32_t count, uint32_t rl,
                                                  uint32_t row_base)
{
    uint32_t i = rl < count ? rl : kNoIndex;
    return row_arg_best_lanes<Max>(v, i, 1, row_base);
}

// AABB (min / max of the transformed points) of hull i with one 16-lane row
// (all 16 lanes call it): each lane transforms points rl, rl + 16, ... and
// keeps their first extremes, then each component is reduced over the row.
// A NaN coordinate, or more than 16 points, sends the hull to the sequential
// scan on the row's first lane.  The bounds are valid in lane rl == 0.
constexpr uint32_t kHullRowLanes = 16;

__device__ __forceinline__ void hull_row_aabb(uint32_t i, const float4* __restrict__ pose,
                                              const float4* __restrict__ shape,
                                              const float* __restrict__ hpts, uint32_t nh,
                                              uint32_t rl, uint32_t row_base, float b[6])
{
    const pg::pose t = load_pose4(pose, i);
    uint32_t first, count;
    hull_range(shape, i, first, count);
    f3 mn{0.0f, 0.0f, 0.0f}, mx{0.0f, 0.0f, 0.0f};
    for (uint32_t k = rl; k < count; k += kHullRowLanes) {
        const uint32_t p = first + k;
        const f3 w = to_world(f3{hpts[p], hpts[nh + p], hpts[2 * nh + p]}, t);
        if (k == rl) {
            mn = w;
            mx = w;
        } else {
            mn = {min_ref(mn.x, w.x), min_ref(mn.y, w.y), min_ref(mn.z, w.z)};
            mx = {max_ref(mx.x, w.x), max_ref(mx.y, w.y), max_ref(mx.z, w.z)};
        }
    }
    // Chunk order within a lane is index order, so a lane's first extreme
    // precedes every later equal value of that lane; across lanes the row
    // reduction takes the lowest lane, i.e. the lowest index among the
    // lanes' first extremes only when count <= 16.  Longer hulls take
    // the sequential scan.
    bool ok = count <= kHullRowLanes;
    b[0] = mn.x;
    b[1] = mn.y;
    b[2] = mn.z;
    b[3] = mx.x;
    b[4] = mx.y;
    b[5] = mx.z;
    if (ok) {
        ok = row_first_extreme<false>(b[0], count, rl, row_base) &&
             row_first_extreme<false>(b[1], count, rl, row_base) &&
             row_first_extreme<false>(b[2], count, rl, row_base) &&
             row_first_extreme<true>(b[3], count, rl, row_base) &&
             row_first_extreme<true>(b[4], count, rl, row_base) &&
             row_first_extreme<true>(b[5], count, rl, row_base);
    }
    if (rl == 0 && !ok) {
        f3 w = to_world(f3{hpts[first], hpts[nh + first], hpts[2 * nh + first]}, t);
        f3 a = w, c = w;
        for (uint32_t k = 1; k < count; ++k) {
            const uint32_t p = first + k;
            w = to_world(f3{hpts[p], hpts[nh + p], hpts[2 * nh + p]}, t);
            a = {min_ref(a.x, w.x), min_ref(a.y, w.y), min_ref(a.z, w.z)};
            c = {max_ref(c.x, w.x), max_ref(c.y, w.y), max_ref(c.z, w.z)};
        }
        b[0] = a.x;
        b[1] = a.y;
        b[2] = a.z;
        b[3] = c.x;
        b[4] = c.y;
        b[5] = c.z;
    }
}

// Hull AABBs (listed by k_hull_list), one 16-lane row per hull.
__global__ __launch_bounds__(kBlock) void k_hull_aabbs(const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ list_count,
                                                       const float4* __restrict__ pose,
                                                       const float4* __restrict__ shape,
                                                       const float* __restrict__ hpts, uint32_t nh,
                                                       const uint32_t* __restrict__ world_of,
                                                       float4* __restrict__ aabb)
{
    const uint32_t hulls = *list_count;
    const uint32_t lane = lane_id(), rl = lane % kHullRowLanes, row_base = lane - rl;
    const uint32_t rows = gridDim.x * blockDim.x / kHullRowLanes;
    const uint32_t rounds = (hulls + rows - 1) / rows;
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t h = r * rows + (blockIdx.x * blockDim.x + threadIdx.x) / kHullRowLanes;
        if (h >= hulls)
            continue;  // whole rows (h is row-uniform)
        const uint32_t i = list[h];
        float b[6];
        hull_row_aabb(i, pose, shape, hpts, nh, rl, row_base, b);
        if (rl == 0)
            store_rec(aabb, i, b, i, world_of ? world_of[i] : 0u);
    }
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
    for (uint32_t off = 32; off > 0; off >>= 1)
        v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
    for (uint32_t off = 32; off > 0; off >>= 1)
        v = min(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (uint32_t off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    return v;
}

// Statistics behind the automatic grid cell: largest x side and largest side
// (float keys), sum of largest sides (2^-16 m fixed point, clamped), centre
// min / max per axis (float keys).  Every reduction is order-independent.
struct WidthStats {
    uint32_t wx = 0, wall = 0;
    uint32_t cmn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, cmx[3] = {0u, 0u, 0u};
    unsigned long long wsum = 0;

    __device__ __forceinline__ void add(const float b[6])
    {
        float w[3];
        for (int a = 0; a < 3; ++a) {
            const float lo = b[a], hi = b[3 + a];
            w[a] = hi - lo;
            // Centre as lo/2 + hi/2 cannot overflow for finite bounds.
            const uint32_t ck = float_key(lo * 0.5f + hi * 0.5f);
            cmn[a] = min(cmn[a], ck);
            cmx[a] = max(cmx[a], ck);
        }
        const float wm = fmaxf(w[0], fmaxf(w[1], w[2]));
        wx = max(wx, float_key(w[0]));
        wall = max(wall, float_key(wm));
        const float fx = fminf(wm, 65536.0f) * 65536.0f;
        wsum += (unsigned long long)fx;
    }

    // Workgroup reduction; thread 0 writes the workgroup's partials.  Every
    // thread of the workgroup calls it.
    __device__ void flush(uint64_t* __restrict__ partials)
    {
        wx = wave_max_u32(wx);
        wall = wave_max_u32(wall);
        wsum = wave_sum_u64(wsum);
        for (int a = 0; a < 3; ++a) {
            cmn[a] = wave_min_u32(cmn[a]);
            cmx[a] = wave_max_u32(cmx[a]);
        }
        __shared__ uint32_t s_u[kBlock / 64][8];
        __shared__ unsigned long long s_sum[kBlock / 64];
        const uint32_t wave = threadIdx.x >> 6;
        if ((threadIdx.x & 63u) == 0) {
            s_u[wave][0] = wx;
            s_u[wave][1] = wall;
            for (int a = 0; a < 3; ++a) {
                s_u[wave][2 + a] = cmn[a];
                s_u[wave][5 + a] = cmx[a];
            }
            s_sum[wave] = wsum;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
                s_u[0][0] = max(s_u[0][0], s_u[w][0]);
                s_u[0][1] = max(s_u[0][1], s_u[w][1]);
                for (int a = 0; a < 3; ++a) {
                    s_u[0][2 + a] = min(s_u[0][2 + a], s_u[w][2 + a]);
                    s_u[0][5 + a] = max(s_u[0][5 + a], s_u[w][5 + a]);
                }
                s_sum[0] += s_sum[w];
            }
            // One partial per value and workgroup (k_fold combines them).
            uint64_t* P = partials + blockIdx.x;
            P[size_t(kSlotWidthMaxX) * kMaxPartialBlocks] = s_u[0][0];
            P[size_t(kSlotWidthMax) * kMaxPartialBlocks] = s_u[0][1];
            P[size_t(kSlotWidthSum) * kMaxPartialBlocks] = s_sum[0];
            for (int a = 0; a < 3; ++a) {
                P[size_t(kSlotCentreMin + a) * kMaxPartialBlocks] = s_u[0][2 + a];
                P[size_t(kSlotCentreMax + a) * kMaxPartialBlocks] = s_u[0][5 + a];
            }
        }
    }
};

__global__ __launch_bounds__(kBlock) void k_reduce(uint32_t n, const float4* __restrict__ aabb,
                                                   uint64_t* __restrict__ partials)
{
    WidthStats st;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        const Rec r = load_rec(aabb, i);
        st.add(r.b);
    }
    st.flush(partials);
}

// One launch for a step: the step counters' reset (workgroup 0; no other
// workgroup touches them), AABBs of every body and the width statistics.
// Workgroups [0, body_groups) take the spheres and boxes (grid-stride), the
// rest take the hulls, one 16-lane row each; every workgroup writes its
// partials.
__global__ __launch_bounds__(kBlock) void k_aabbs_step(
    uint32_t n, const float4* __restrict__ pose, const float4* __restrict__ shape,
    const float* __restrict__ hpts, uint32_t nh, const uint32_t* __restrict__ world_of,
    const uint32_t* __restrict__ hull_list, const uint32_t* __restrict__ hull_count,
    uint32_t body_groups, float4* __restrict__ aabb, uint64_t* __restrict__ partials,
    Counters* counters)
{
    if (blockIdx.x == 0)
        reset_counters(counters);
    WidthStats st;
    if (blockIdx.x < body_groups) {
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += body_groups * blockDim.x) {
            float b[6];
            if (body_aabb(i, pose, shape, b)) {
                store_rec(aabb, i, b, i, world_of ? world_of[i] : 0u);
                st.add(b);
            }
        }
    } else {
        const uint32_t hulls = *hull_count;
        const uint32_t lane = lane_id(), rl = lane % kHullRowLanes, row_base = lane - rl;
        const uint32_t rows = (gridDim.x - body_groups) * blockDim.x / kHullRowLanes;
        const uint32_t row = ((blockIdx.x - body_groups) * blockDim.x + threadIdx.x) / kHullRowLanes;
        for (uint32_t h = row; h < hulls; h += rows) {
            const uint32_t i = hull_list[h];
            float b[6];
            hull_row_aabb(i, pose, shape, hpts, nh, rl, row_base, b);
            if (rl == 0) {
                store_rec(aabb, i, b, i, world_of ? world_of[i] : 0u);
                st.add(b);
            }
        }
    }
    st.flush(partials);
}

// The collision step's AABB refresh fused with the uniform grid's keys: one
// workgroup of kKeyThreads per radix-sort tile of kSortThreads * kItems
// colliders.  Spheres and boxes take one thread each; the tile's hulls are
// listed in LDS and take one 16-lane row each (64 rows per round).  Every
// collider's record is written, then its cell key (grid_cell_key with the
// cell from the shapes, shape_cell), its sort item (key | index << 32), its
// in-grid flag and, for an overflow body, a slot of the overflow list; the
// tile's histogram of the bucket digit (key bits low .. low + 7, in-grid
// bodies) goes to hist[digit * tiles + tile], the layout the radix pass
// reads.  Workgroup 0 also resets the step counters (no other workgroup
// touches them: the overflow slots come from over_ctr).
constexpr uint32_t kKeyThreads = 1024;

template <uint32_t kItems>
__global__ __launch_bounds__(kKeyThreads) void k_aabbs_grid_keys(
    uint32_t n, const float4* __restrict__ pose, const float4* __restrict__ shape,
    const float* __restrict__ hpts, uint32_t nh, const uint32_t* __restrict__ world_of,
    float4* __restrict__ aabb, Counters* counters, const uint32_t* __restrict__ shape_cell,
    float requested, uint32_t bits_per_axis, uint32_t table_bits, uint32_t low,
    uint64_t* __restrict__ items, uint32_t* __restrict__ overflow,
    uint32_t* __restrict__ over_ctr, uint32_t* __restrict__ hist, uint32_t* __restrict__ bounds,
    uint32_t bound_words, uint64_t* __restrict__ epa_partials, uint32_t* __restrict__ epa_epoch)
{
    constexpr uint32_t kTile = kSortThreads * kItems;
    // The narrow phase of this step: a new epoch for its EPA hand-over flags,
    // and its EPA partial slots zeroed (the work of its classification, which
    // the grid query does).
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *epa_epoch += 1u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kEpaPartialSlots * kEpaPartials;
         i += gridDim.x * blockDim.x)
        epa_partials[(i / kEpaPartials) * kMaxPartialBlocks + i % kEpaPartials] = 0;
    // Bucket ranges of a two-pass sort start empty.
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < bound_words;
         i += gridDim.x * blockDim.x)
        bounds[i] = 0;
    static_assert(kTile % kKeyThreads == 0, "whole rounds per tile");
    __shared__ uint32_t counts[256];
    __shared__ uint32_t hull_ids[kTile];
    __shared__ uint32_t hull_count;
    if (blockIdx.x == 0)
        reset_counters(counters);
    const float cell = requested > 0.0f ? requested : __uint_as_float(*shape_cell);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        counters->cell_bits = __float_as_uint(cell);
    if (threadIdx.x < 256)
        counts[threadIdx.x] = 0;
    if (threadIdx.x == 0)
        hull_count = 0;
    __syncthreads();
    const double inv = 1.0 / double(cell);
    auto emit = [&](uint32_t i, const float b[6], uint32_t world) {
        const uint32_t key = grid_cell_key(b, world, cell, inv, bits_per_axis, table_bits);
        items[i] = uint64_t(key) | (uint64_t(i) << 32);
        if (key == kOverflowKey)
            overflow[atomicAdd(over_ctr, 1u)] = i;
        else
            atomicAdd(&counts[(key >> low) & 0xFFu], 1u);
    };
    const uint32_t base = blockIdx.x * kTile;
#pragma unroll
    for (uint32_t r = 0; r < kTile / kKeyThreads; ++r) {
        const uint32_t i = base + r * kKeyThreads + threadIdx.x;
        if (i < n) {
            float b[6];
            if (body_aabb(i, pose, shape, b)) {
                const uint32_t w = world_of ? world_of[i] : 0u;
                const uint32_t ty = __float_as_uint(shape[2 * size_t(i)].w);
                store_rec(aabb, i, b, i, w | (ty << kRecTypeShift));
                emit(i, b, w);
            } else {
                hull_ids[atomicAdd(&hull_count, 1u)] = i;
            }
        }
    }
    __syncthreads();
    const uint32_t hulls = hull_count;
    const uint32_t lane = lane_id(), rl = lane % kHullRowLanes, row_base = lane - rl;
    constexpr uint32_t kRows = kKeyThreads / kHullRowLanes;
    const uint32_t row = threadIdx.x / kHullRowLanes;
    const uint32_t rounds = (hulls + kRows - 1) / kRows;
    for (uint32_t k = 0; k < rounds; ++k) {
        const uint32_t h = k * kRows + row;
        if (h >= hulls)
            continue;  // whole rows (h is row-uniform)
        const uint32_t i = hull_ids[h];
        float b[6];
        hull_row_aabb(i, pose, shape, hpts, nh, rl, row_base, b);
        if (rl == 0) {
            const uint32_t w = world_of ? world_of[i] : 0u;
            store_rec(aabb, i, b, i, w | (uint32_t(PHYSGPU_HULL) << kRecTypeShift));
            emit(i, b, w);
        }
    }
    __syncthreads();
    const uint32_t tiles = (n + kTile - 1) / kTile;
    if (threadIdx.x < 256)
        hist[threadIdx.x * tiles + blockIdx.x] = counts[threadIdx.x];
}

// Largest AABB side any orientation can give each collider (sphere 2 r, box
// 2 |h|, hull 2 max |p_k|): their maximum (float key, out[0]) and sum
// (2^-16 m fixed point, out[1..2]).  One atomic per workgroup; run once per
// upload.
__global__ __launch_bounds__(kBlock) void k_shape_bound(uint32_t n, const float4* __restrict__ shape,
                                                        const float* __restrict__ hpts, uint32_t nh,
                                                        uint32_t* __restrict__ out)
{
    uint32_t mx = 0;
    unsigned long long sum = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 s0 = shape[2 * size_t(i)];
        const uint32_t ty = __float_as_uint(s0.w);
        float d;
        if (ty == PHYSGPU_SPHERE) {
            d = 2.0f * s0.x;
        } else if (ty == PHYSGPU_BOX) {
            d = 2.0f * sqrtf(s0.x * s0.x + s0.y * s0.y + s0.z * s0.z);
        } else {
            uint32_t first, count;
            hull_range(shape, i, first, count);
            float r2 = 0.0f;
            for (uint32_t k = 0; k < count; ++k) {
                const uint32_t p = first + k;
                const float x = hpts[p], y = hpts[nh + p], z = hpts[2 * nh + p];
                r2 = fmaxf(r2, x * x + y * y + z * z);
            }
            d = 2.0f * sqrtf(r2);
        }
        if (!(d >= 0.0f))
            d = 0.0f;  // NaN data: such a body takes the overflow path anyway
        mx = max(mx, float_key(d));
        sum += (unsigned long long)(fminf(d, 65536.0f) * 65536.0f);
    }
    mx = wave_max_u32(mx);
    sum = wave_sum_u64(sum);
    if (lane_id() == 0) {
        atomicMax(&out[0], mx);
        atomicAdd(reinterpret_cast<unsigned long long*>(out + 2), sum);
    }
}

// Grid cell of the collision step from the shape bounds: the largest bound,
// capped at 8x the mean bound (larger bodies take the overflow path), widened
// by 2^-8 (in-grid bodies must be at most 1 - 2^-9 of the cell); out[1].
__global__ void k_shape_cell(uint32_t n, uint32_t* __restrict__ out)
{
    const float dmax = key_float(out[0]);
    const unsigned long long sum = *reinterpret_cast<const unsigned long long*>(out + 2);
    const double mean = n ? double(sum) / 65536.0 / double(n) : 1.0;
    float cell = fminf(dmax, float(8.0 * mean));
    if (!(cell > 0.0f))
        cell = 1.0f;
    out[1] = __float_as_uint(cell * (1.0f + 1.0f / 256.0f));
}

__global__ __launch_bounds__(kBlock) void k_rec_to_aabb(uint32_t n, const float4* __restrict__ rec,
                                                        physgpu_aabb* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const Rec r = load_rec(rec, i);
    physgpu_aabb b;
    for (int a = 0; a < 3; ++a) {
        b.min[a] = r.b[a];
        b.max[a] = r.b[3 + a];
    }
    out[i] = b;
}

__global__ __launch_bounds__(kBlock) void k_aabb_to_rec(uint32_t n, const physgpu_aabb* __restrict__ in,
                                                        float4* __restrict__ rec)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const physgpu_aabb b = in[i];
    const float f[6] = {b.min[0], b.min[1], b.min[2], b.max[0], b.max[1], b.max[2]};
    store_rec(rec, i, f, i);
}

// World pose of each collider: body pose composed with the collider's local
// transform, as physics_world.cpp:660-667 and aabb.cpp:10-13 compose it.
__global__ __launch_bounds__(kBlock) void k_compose(uint32_t n, const float* __restrict__ body,
                                                    const float* __restrict__ local,
                                                    float4* __restrict__ pose_out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const f3 bp{body[i], body[n + i], body[2 * n + i]};
    const quat bq{body[3 * n + i], body[4 * n + i], body[5 * n + i], body[6 * n + i]};
    f3 p = bp;
    quat q = bq;
    if (local) {
        const f3 lp{local[i], local[n + i], local[2 * n + i]};
        const quat lq{local[3 * n + i], local[4 * n + i], local[5 * n + i], local[6 * n + i]};
        p = add(bp, rotate(bq, lp));
        q = qmul(bq, lq);
    }
    store_pose4(pose_out, i, pg::pose{p, q});
}

// Resident layouts from the uploaded planes: poses (7 x n) to 2 float4 per
// body; types, shape data (3 x n) and hull ranges (2 x n) to 2 float4 per
// body (see load_pose4 / shape_data).
__global__ __launch_bounds__(kBlock) void k_pose_aos(uint32_t n, const float* __restrict__ p,
                                                     float4* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    store_pose4(out, i, pg::pose{{p[i], p[n + i], p[2 * n + i]},
                                 {p[3 * n + i], p[4 * n + i], p[5 * n + i], p[6 * n + i]}});
}

__global__ __launch_bounds__(kBlock) void k_shape_aos(uint32_t n, const uint8_t* __restrict__ type,
                                                      const float* __restrict__ sdata,
                                                      const uint32_t* __restrict__ hrange,
                                                      float4* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t ty = type[i];
    const uint32_t first = ty == PHYSGPU_HULL ? hrange[i] : 0u;
    const uint32_t count = ty == PHYSGPU_HULL ? hrange[n + i] : 0u;
    out[2 * size_t(i)] = make_float4(sdata[i], sdata[n + i], sdata[2 * n + i], __uint_as_float(ty));
    out[2 * size_t(i) + 1] = make_float4(__uint_as_float(first), __uint_as_float(count), 0.0f, 0.0f);
}

}  // namespace

void launch_compose_world(physgpu_context* ctx)
{
    if (ctx->n == 0)
        return;
    Timer t;
    time_begin(ctx, "compose_world", t);
    hipLaunchKernelGGL(k_compose, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->n, ctx->world_body_pose,
                       ctx->world_has_local ? ctx->local_pose : nullptr, ctx->pose4);
    time_end(ctx, t);
}

void launch_step_begin(physgpu_context* ctx)
{
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, ctx->stream, ctx->counters);
}

namespace {

void fold_width_stats(physgpu_context* ctx, uint32_t groups)
{
    FoldList fold{};
    fold.e[0] = {kSlotWidthMaxX, groups, kFoldMax32, uint32_t(offsetof(Counters, wmax_x_key))};
    fold.e[1] = {kSlotWidthMax, groups, kFoldMax32, uint32_t(offsetof(Counters, wmax_key))};
    fold.e[2] = {kSlotWidthSum, groups, kFoldSum64, uint32_t(offsetof(Counters, width_sum_fx))};
    for (uint32_t a = 0; a < 3; ++a) {
        fold.e[3 + a] = {kSlotCentreMin + a, groups, kFoldMin32,
                         uint32_t(offsetof(Counters, cmin_key) + 4 * a)};
        fold.e[6 + a] = {kSlotCentreMax + a, groups, kFoldMax32,
                         uint32_t(offsetof(Counters, cmax_key) + 4 * a)};
    }
    fold.count = 9;
    launch_fold(ctx, fold);
}

}  // namespace

void launch_compute_aabbs(physgpu_context* ctx, bool with_stats)
{
    ctx->resident_aabbs_valid = true;
    if (ctx->n == 0) {
        if (with_stats)
            launch_step_begin(ctx);
        return;
    }
    Timer t;
    time_begin(ctx, "aabb_from_pose", t);
    const uint32_t* world_of = ctx->n_worlds ? ctx->world_of : nullptr;
    if (with_stats) {
        // Spheres and boxes: up to 2,048 workgroups; hulls: the rest of the
        // partial slots, one row per hull and round.
        const uint32_t body_groups = std::min<uint32_t>(div_up(ctx->n, kBlock), 2048);
        const uint32_t hull_groups =
            ctx->has_hulls ? std::min<uint32_t>(div_up(16ull * ctx->n, kBlock),
                                                kMaxPartialBlocks - body_groups)
                           : 0u;
        hipLaunchKernelGGL(k_aabbs_step, dim3(body_groups + hull_groups), dim3(kBlock), 0,
                           ctx->stream, ctx->n, ctx->pose4, ctx->shape4, ctx->hull_points,
                           ctx->n_hull_points, world_of,
                           ctx->hull_list, ctx->hull_list_count, body_groups, ctx->aabb,
                           ctx->partials, ctx->counters);
        fold_width_stats(ctx, body_groups + hull_groups);
        time_end(ctx, t);
        return;
    }
    hipLaunchKernelGGL(k_compute_aabbs, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0,
                       ctx->stream, ctx->n, ctx->pose4, ctx->shape4, world_of, ctx->aabb);
    if (ctx->has_hulls)
        hipLaunchKernelGGL(k_hull_aabbs, dim3(stride_grid(16ull * ctx->n)), dim3(kBlock), 0,
                           ctx->stream, ctx->hull_list, ctx->hull_list_count, ctx->pose4,
                           ctx->shape4, ctx->hull_points, ctx->n_hull_points, world_of, ctx->aabb);
    time_end(ctx, t);
}

void launch_shape_cell(physgpu_context* ctx)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->shape_cell, 4u, 0u);
    if (ctx->n)
        hipLaunchKernelGGL(k_shape_bound, dim3(stride_grid(ctx->n)), dim3(kBlock), 0, ctx->stream,
                           ctx->n, ctx->shape4, ctx->hull_points, ctx->n_hull_points,
                           ctx->shape_cell);
    hipLaunchKernelGGL(k_shape_cell, dim3(1), dim3(1), 0, ctx->stream, ctx->n, ctx->shape_cell);
}

void launch_aabbs_grid_keys(physgpu_context* ctx, float cell, const GridBits& g)
{
    ctx->resident_aabbs_valid = true;
    Timer t;
    time_begin(ctx, "aabb_grid_keys", t);
    const uint32_t* world_of = ctx->n_worlds ? ctx->world_of : nullptr;
    const uint32_t tile = kSortThreads * ctx->tuning.sort_items;
    const dim3 grid(div_up(ctx->n, tile)), block(kKeyThreads);
#define PG_AABB_KEYS(K)                                                                        \
    hipLaunchKernelGGL(k_aabbs_grid_keys<K>, grid, block, 0, ctx->stream, ctx->n, ctx->pose4,   \
                       ctx->shape4, ctx->hull_points, ctx->n_hull_points, world_of, ctx->aabb, \
                       ctx->counters, ctx->shape_cell + 1, cell, g.axis, g.table, g.low,       \
                       ctx->items_a, ctx->overflow_list, ctx->grid_over_ctr, ctx->histogram,   \
                       ctx->bucket_bounds, g.bucket > 8 ? 2u << g.bucket : 0u,                  \
                       ctx->partials + size_t(kEpaPartialFirst) * kMaxPartialBlocks, ctx->epa_epoch)
    if (ctx->tuning.sort_items == 4)
        PG_AABB_KEYS(4);
    else if (ctx->tuning.sort_items == 16)
        PG_AABB_KEYS(16);
    else
        PG_AABB_KEYS(8);
#undef PG_AABB_KEYS
    time_end(ctx, t);
}

void launch_pose_aos(physgpu_context* ctx, const float* src)
{
    if (ctx->n)
        hipLaunchKernelGGL(k_pose_aos, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0, ctx->stream,
                           ctx->n, src ? src : ctx->pose, ctx->pose4);
}

void launch_shape_aos(physgpu_context* ctx)
{
    if (ctx->n)
        hipLaunchKernelGGL(k_shape_aos, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0, ctx->stream,
                           ctx->n, ctx->shape_type, ctx->shape_data, ctx->hull_range, ctx->shape4);
}

void launch_hull_list(physgpu_context* ctx)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->hull_list_count, 1u,
                       0u);
    if (ctx->n)
        hipLaunchKernelGGL(k_hull_list, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0,
                           ctx->stream, ctx->n, ctx->shape_type, ctx->hull_list,
                           ctx->hull_list_count);
}

void launch_aabb_reductions(physgpu_context* ctx)
{
    if (ctx->n == 0)
        return;
    Timer t;
    time_begin(ctx, "aabb_reduce", t);
    const uint32_t grid = std::min<uint32_t>(div_up(ctx->n, kBlock), 1024);
    hipLaunchKernelGGL(k_reduce, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->aabb,
                       ctx->partials);
    fold_width_stats(ctx, grid);
    time_end(ctx, t);
}

void launch_convert_aabbs_out(physgpu_context* ctx, physgpu_aabb* dev_out)
{
    if (ctx->n == 0)
        return;
    hipLaunchKernelGGL(k_rec_to_aabb, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->n, ctx->aabb, dev_out);
}

void launch_convert_aabbs_in(physgpu_context* ctx, const physgpu_aabb* dev_in)
{
    if (ctx->n == 0)
        return;
    hipLaunchKernelGGL(k_aabb_to_rec, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->n, dev_in, ctx->aabb);
}

}  // namespace pg
