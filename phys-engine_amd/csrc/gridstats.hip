// The reference grid's statistics (findGridPairs, src/collision/broadphase.cpp:90-172),
// computed on demand from the resident AABB records.  The GPU's own grid
// (grid.hip) finds the same pairs differently (one cell per body, cell = the
// largest side); these are the numbers the reference's grid reports:
//   gridCellSize      upper median of the positive finite widths (largest side,
//                     in double), 1.0 when there is none (:94-110);
//   gridOverflowAabbs AABBs whose cell range fails gridRange (:62-88: not
//                     finite, outside int32, or more than 64 cells);
//   gridEntries       cells of the other AABBs, one entry per cell (:124-128);
//   gridComparisons   entry pairs sharing a cell, the sum over cells of C(k, 2)
//                     (:140-150), plus the overflow pass (:152-155): each
//                     overflow AABB against every in-grid AABB and every later
//                     overflow AABB, O G + C(O, 2).
// Widths are sorted with the radix sort (positive doubles order as their bit
// patterns).  Entries are counted per cell in an open-addressing table keyed
// by a 64-bit hash of the cell; each slot's smallest entry id is its owner,
// and an entry whose cell differs from its owner's (a hash collision) is
// counted in a further round with another hash, so every cell is counted
// exactly once, in one slot.  PHYSGPU_GRID_STATS_TAG_BITS (1..64, default 64)
// keeps only that many low bits of each tag: a test setting forces hash
// collisions, hence further rounds, with results unchanged.
#include <stddef.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "wave.hpp"

namespace pg {
namespace {

constexpr uint64_t kNoWidth = ~0ull;
constexpr uint32_t kMaxRounds = 8;

struct GridScalars {
    unsigned long long valid;     // positive finite widths
    unsigned long long entries;   // cells of in-grid AABBs
    unsigned long long overflow;  // AABBs failing gridRange
    unsigned long long shared;    // sum over cells of C(k, 2)
    unsigned int collisions;      // entries left for the next round
    unsigned int pad;
    double cell;
};

// std::max({x, y, z}): the first of equal maxima, by operator<.
__device__ __forceinline__ double width_of(const Rec& r)
{
    double w = double(r.b[3]) - double(r.b[0]);
    const double y = double(r.b[4]) - double(r.b[1]), z = double(r.b[5]) - double(r.b[2]);
    if (w < y)
        w = y;
    if (w < z)
        w = z;
    return w;
}

__global__ void k_gs_init(GridScalars* gs, uint32_t count)
{
    for (uint32_t k = 0; k < count; ++k) {
        gs[k].valid = gs[k].entries = gs[k].overflow = gs[k].shared = 0;
        gs[k].collisions = gs[k].pad = 0;
        gs[k].cell = 1.0;
    }
}

__global__ __launch_bounds__(kBlock) void k_gs_widths(uint32_t n, const float4* __restrict__ aabb,
                                                      uint64_t* __restrict__ keys,
                                                      GridScalars* gs)
{
    unsigned long long valid = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        const double w = width_of(load_rec(aabb, i));
        const bool ok = w > 0.0 && isfinite(w);
        keys[i] = ok ? uint64_t(__double_as_longlong(w)) : kNoWidth;
        valid += ok ? 1u : 0u;
    }
    block_add(valid, &gs->valid);
}

// nth_element(begin + size / 2) of the valid widths: position valid / 2 of
// the sorted keys (the others sort last).
__global__ void k_gs_cell(const uint64_t* __restrict__ sorted, GridScalars* gs)
{
    const unsigned long long m = gs->valid;
    gs->cell = m ? __longlong_as_double((long long)sorted[m / 2]) : 1.0;
}

// gridRange (broadphase.cpp:62-88): true with the first cell and the count
// per axis, false for the overflow path.
__device__ __forceinline__ bool grid_range(const Rec& r, double cell, int32_t first[3],
                                           int32_t count[3])
{
    uint64_t total = 1;
    for (int a = 0; a < 3; ++a) {
        const double lower = floor(double(r.b[a]) / cell);
        const double upper = floor(double(r.b[3 + a]) / cell);
        if (!isfinite(lower) || !isfinite(upper) || lower < -2147483648.0 ||
            upper > 2147483647.0 || lower > upper)
            return false;
        const int64_t c = int64_t(upper) - int64_t(lower) + 1;
        if (c > int64_t(64 / total))
            return false;
        first[a] = int32_t(lower);
        count[a] = int32_t(c);
        total *= uint64_t(c);
    }
    return true;
}

// Per record: its first cell and the counts per axis packed in w (x | y << 8
// | z << 16; 0 on the overflow path); entry and overflow totals.
__global__ __launch_bounds__(kBlock) void k_gs_ranges(uint32_t n, const float4* __restrict__ aabb,
                                                      int4* __restrict__ ranges, GridScalars* gs)
{
    const double cell = gs->cell;
    unsigned long long entries = 0, overflow = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        int32_t f[3] = {0, 0, 0}, c[3] = {0, 0, 0};
        if (grid_range(load_rec(aabb, i), cell, f, c)) {
            entries += uint64_t(c[0] * c[1] * c[2]);
            ranges[i] = make_int4(f[0], f[1], f[2], c[0] | (c[1] << 8) | (c[2] << 16));
        } else {
            ++overflow;
            ranges[i] = make_int4(0, 0, 0, 0);
        }
    }
    block_add(entries, &gs->entries);
    block_add(overflow, &gs->overflow);
}

// Entry id: record index * 64 + the cell's rank in the record's x-major
// range (at most 64 cells).
struct Cell {
    int32_t x, y, z;
};

__device__ __forceinline__ Cell cell_of(const int4* __restrict__ ranges, uint64_t eid)
{
    const int4 r = ranges[eid >> 6];
    const uint32_t l = uint32_t(eid & 63u);
    const uint32_t cy = (uint32_t(r.w) >> 8) & 0xFFu, cz = (uint32_t(r.w) >> 16) & 0xFFu;
    return {r.x + int32_t(l / (cy * cz)), r.y + int32_t((l / cz) % cy), r.z + int32_t(l % cz)};
}

__device__ __forceinline__ uint64_t mix64(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// Nonzero tag of a cell under the round's seed (0 marks an empty slot).
__device__ __forceinline__ uint64_t cell_tag(Cell c, uint32_t seed, uint64_t tag_mask)
{
    const uint64_t xy = uint64_t(uint32_t(c.x)) << 32 | uint32_t(c.y);
    return (mix64(mix64(xy ^ (uint64_t(seed) * 0x9e3779b97f4a7c15ull)) ^ uint32_t(c.z)) &
            tag_mask) |
           1ull;
}

// Start of a tag's probe sequence (multiplicative hashing, so that tags kept
// to a few low bits still spread over the table).
__device__ __forceinline__ uint64_t first_slot(uint64_t tag, uint64_t mask)
{
    return ((tag * 0x9e3779b97f4a7c15ull) >> 20) & mask;
}

// Entries of a round: every cell of every record (list == nullptr), or the
// ids in list[0, prev->collisions).
template <typename F>
__device__ __forceinline__ void for_each_entry(uint32_t n, const int4* __restrict__ ranges,
                                               const uint64_t* __restrict__ list,
                                               const GridScalars* prev, F&& f)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    if (!list) {
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            const uint32_t w = uint32_t(ranges[i].w);
            const uint32_t cells = (w & 0xFFu) * ((w >> 8) & 0xFFu) * ((w >> 16) & 0xFFu);
            for (uint32_t l = 0; l < cells; ++l)
                f(uint64_t(i) * 64u + l);
        }
    } else {
        const uint32_t m = prev->collisions;
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += stride)
            f(list[k]);
    }
}

// Claim: the first slot on the probe sequence that is empty or holds the
// tag; the smallest entry id landing there owns the slot.  The table holds
// at least twice the round's entries, so every probe finds a slot.
__global__ __launch_bounds__(kBlock) void k_gs_claim(uint32_t n, const int4* __restrict__ ranges,
                                                     const uint64_t* __restrict__ list,
                                                     const GridScalars* prev, uint32_t seed,
                                                     uint64_t tag_mask,
                                                     uint64_t* __restrict__ tags,
                                                     uint64_t* __restrict__ owner, uint64_t mask)
{
    for_each_entry(n, ranges, list, prev, [&](uint64_t eid) {
        const uint64_t tag = cell_tag(cell_of(ranges, eid), seed, tag_mask);
        uint64_t s = first_slot(tag, mask);
        for (;;) {
            const unsigned long long was =
                atomicCAS(reinterpret_cast<unsigned long long*>(tags + s), 0ull,
                          (unsigned long long)tag);
            if (was == 0ull || was == tag) {
                atomicMin(reinterpret_cast<unsigned long long*>(owner + s),
                          (unsigned long long)eid);
                return;
            }
            s = (s + 1) & mask;
        }
    });
}

// Count: tags no longer change, so the first slot holding the entry's tag is
// the one it claimed.  Same cell as the owner: counted there; otherwise a
// hash collision, listed for the next round.
__global__ __launch_bounds__(kBlock) void k_gs_count(uint32_t n, const int4* __restrict__ ranges,
                                                     const uint64_t* __restrict__ list,
                                                     const GridScalars* prev, uint32_t seed,
                                                     uint64_t tag_mask,
                                                     const uint64_t* __restrict__ tags,
                                                     const uint64_t* __restrict__ owner,
                                                     uint32_t* __restrict__ counts, uint64_t mask,
                                                     uint64_t* __restrict__ next,
                                                     GridScalars* gs)
{
    for_each_entry(n, ranges, list, prev, [&](uint64_t eid) {
        const Cell c = cell_of(ranges, eid);
        const uint64_t tag = cell_tag(c, seed, tag_mask);
        uint64_t s = first_slot(tag, mask);
        while (tags[s] != tag)
            s = (s + 1) & mask;
        const Cell o = cell_of(ranges, owner[s]);
        if (o.x == c.x && o.y == c.y && o.z == c.z)
            atomicAdd(counts + s, 1u);
        else
            next[atomicAdd(&gs->collisions, 1u)] = eid;
    });
}

__global__ __launch_bounds__(kBlock) void k_gs_pairs(const uint32_t* __restrict__ counts,
                                                     uint64_t size, GridScalars* gs)
{
    unsigned long long shared = 0;
    for (uint64_t h = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; h < size;
         h += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t k = counts[h];
        shared += k ? k * (k - 1) / 2 : 0;
    }
    block_add(shared, &gs->shared);
}

__global__ __launch_bounds__(kBlock) void k_gs_clear(uint64_t* __restrict__ tags,
                                                     uint64_t* __restrict__ owner,
                                                     uint32_t* __restrict__ counts, uint64_t size)
{
    for (uint64_t h = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; h < size;
         h += uint64_t(gridDim.x) * blockDim.x) {
        tags[h] = 0ull;
        owner[h] = ~0ull;
        counts[h] = 0u;
    }
}

// The round's own collision counter starts at zero (its other block holds
// the previous round's, which sizes this round's list).
__global__ void k_gs_next_round(GridScalars* cur) { cur->collisions = 0; }

template <typename T>
hipError_t grow(T*& p, size_t& have, size_t need)
{
    if (have >= need)
        return hipSuccess;
    if (p)
        (void)hipFree(p);
    p = nullptr;
    have = 0;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), need * sizeof(T));
    if (e == hipSuccess)
        have = need;
    return e;
}

struct WorldGridStats {
    double cell;
    uint64_t entries, overflow, comparisons;
};

// The reference grid's statistics of records [0, n) of `aabb` (one world).
physgpu_status world_grid_stats(physgpu_context* ctx, const float4* aabb, uint32_t n,
                                uint64_t tag_mask, WorldGridStats& out)
{
    out = {1.0, 0, 0, 0};
    if (n == 0)
        return PHYSGPU_OK;
    hipStream_t st = ctx->stream;
    // Two scalar blocks, used by alternate hash rounds: each round counts its
    // collisions in its own block and reads its list length from the other.
    GridScalars* gs = reinterpret_cast<GridScalars*>(ctx->gs_scalars);
    GridScalars* gs2 = gs + 1;
    const uint32_t grid = stride_grid(n);
    hipLaunchKernelGGL(k_gs_init, dim3(1), dim3(1), 0, st, gs, 2u);
    hipLaunchKernelGGL(k_gs_widths, dim3(grid), dim3(kBlock), 0, st, n, aabb, ctx->items_a, gs);
    uint64_t* sorted = nullptr;
    radix_sort(ctx, ctx->items_a, ctx->items_b, n, nullptr, 0, 32, 32, 32, &sorted);
    hipLaunchKernelGGL(k_gs_cell, dim3(1), dim3(1), 0, st, sorted, gs);
    hipLaunchKernelGGL(k_gs_ranges, dim3(grid), dim3(kBlock), 0, st, n, aabb, ctx->gs_ranges,
                       gs);
    GridScalars h;
    PG_TRY_HIP(hipGetLastError());
    PG_TRY_HIP(hipMemcpyAsync(&h, gs, sizeof(h), hipMemcpyDeviceToHost, st));
    PG_TRY_HIP(hipStreamSynchronize(st));
    out.cell = h.cell;
    out.entries = h.entries;
    out.overflow = h.overflow;
    const uint64_t o = h.overflow, g = n - h.overflow;
    out.comparisons = o * g + (o ? o * (o - 1) / 2 : 0);
    if (h.entries == 0)
        return PHYSGPU_OK;

    uint64_t size = 1024;
    while (size < 2 * h.entries)
        size <<= 1;
    PG_TRY_HIP(grow(ctx->gs_tags, ctx->gs_tags_cap, size));
    PG_TRY_HIP(grow(ctx->gs_owner, ctx->gs_owner_cap, size));
    PG_TRY_HIP(grow(ctx->gs_counts, ctx->gs_counts_cap, size));
    PG_TRY_HIP(grow(ctx->gs_list, ctx->gs_list_cap, 2 * h.entries));
    uint64_t* lists[2] = {ctx->gs_list, ctx->gs_list + h.entries};
    // Round r reads the list round r - 1 wrote (none for round 0) and writes
    // the other one; a round's entries are at most the previous round's.
    uint64_t active = h.entries;
    for (uint32_t round = 0;; ++round) {
        if (round == kMaxRounds)
            return fail(PHYSGPU_DEVICE_ERROR, "grid statistics: cell hash collisions persist");
        uint64_t rsize = 1024;
        while (rsize < 2 * active)
            rsize <<= 1;
        GridScalars* cur = (round & 1) ? gs2 : gs;
        GridScalars* prev = (round & 1) ? gs : gs2;
        const uint64_t* list = round ? lists[(round - 1) & 1] : nullptr;
        uint64_t* next = lists[round & 1];
        const uint32_t tgrid =
            uint32_t(std::min<uint64_t>(div_up(rsize, uint64_t(kBlock)), kStrideGrid));
        const uint32_t egrid = round ? stride_grid(active) : grid;
        hipLaunchKernelGGL(k_gs_clear, dim3(tgrid), dim3(kBlock), 0, st, ctx->gs_tags,
                           ctx->gs_owner, ctx->gs_counts, rsize);
        hipLaunchKernelGGL(k_gs_next_round, dim3(1), dim3(1), 0, st, cur);
        hipLaunchKernelGGL(k_gs_claim, dim3(egrid), dim3(kBlock), 0, st, n, ctx->gs_ranges, list,
                           prev, round, tag_mask, ctx->gs_tags, ctx->gs_owner, rsize - 1);
        hipLaunchKernelGGL(k_gs_count, dim3(egrid), dim3(kBlock), 0, st, n, ctx->gs_ranges, list,
                           prev, round, tag_mask, ctx->gs_tags, ctx->gs_owner, ctx->gs_counts,
                           rsize - 1,
                           next, cur);
        hipLaunchKernelGGL(k_gs_pairs, dim3(tgrid), dim3(kBlock), 0, st, ctx->gs_counts, rsize,
                           cur);
        PG_TRY_HIP(hipGetLastError());
        PG_TRY_HIP(hipMemcpyAsync(&h, cur, sizeof(h), hipMemcpyDeviceToHost, st));
        PG_TRY_HIP(hipStreamSynchronize(st));
        if (h.collisions == 0)
            break;
        active = h.collisions;
    }
    // The cell pair sums of every round, accumulated in the two blocks.
    GridScalars h2;
    PG_TRY_HIP(hipMemcpyAsync(&h, gs, sizeof(h), hipMemcpyDeviceToHost, st));
    PG_TRY_HIP(hipMemcpyAsync(&h2, gs2, sizeof(h2), hipMemcpyDeviceToHost, st));
    PG_TRY_HIP(hipStreamSynchronize(st));
    out.comparisons += h.shared + h2.shared;
    return PHYSGPU_OK;
}

}  // namespace

physgpu_status reference_grid_stats(physgpu_context* ctx, physgpu_broadphase_stats* s)
{
    s->grid_cell_size = 1.0;
    s->grid_entries = s->grid_comparisons = s->grid_overflow_aabbs = 0;
    if (ctx->n == 0)
        return PHYSGPU_OK;
    if (!ctx->resident_aabbs_valid)
        return fail(PHYSGPU_INVALID_ARGUMENT,
                    "grid statistics: a proxy-tree query replaced the resident AABBs");
    // The width sort reuses the broad phase's item buffers: a kept LBVH build
    // is no longer current.
    ctx->tree_build_current = false;
    PG_TRY_HIP(grow(ctx->gs_scalars, ctx->gs_scalars_cap, 2 * sizeof(GridScalars)));
    PG_TRY_HIP(grow(ctx->gs_ranges, ctx->gs_ranges_cap, ctx->n));
    uint64_t tag_mask = ~0ull;
    if (const char* bits = std::getenv("PHYSGPU_GRID_STATS_TAG_BITS")) {
        const long b = std::strtol(bits, nullptr, 10);
        if (b >= 1 && b < 64)
            tag_mask = (1ull << b) - 1;
    }
    std::vector<uint32_t> off;
    if (ctx->n_worlds)
        off = ctx->host_world_offsets;
    else
        off = {0u, ctx->n};
    // A batch of worlds: each world is its own reference grid; the counts add
    // up and the cell reported is the largest world cell.
    double cell = 0.0;
    for (size_t w = 0; w + 1 < off.size(); ++w) {
        if (off[w + 1] == off[w])
            continue;
        WorldGridStats ws;
        const physgpu_status st =
            world_grid_stats(ctx, ctx->aabb + 2 * size_t(off[w]), off[w + 1] - off[w], tag_mask, ws);
        if (st != PHYSGPU_OK)
            return st;
        cell = std::max(cell, ws.cell);
        s->grid_entries += ws.entries;
        s->grid_overflow_aabbs += ws.overflow;
        s->grid_comparisons += ws.comparisons;
    }
    s->grid_cell_size = cell > 0.0 ? cell : 1.0;
    return PHYSGPU_OK;
}

}  // namespace pg
