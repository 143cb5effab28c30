// Uniform-grid broad phase (same pair set as findGridPairs,
// src/collision/broadphase.cpp:90-182: every inclusive AABB overlap, once).
//
// The reference inserts each AABB into up to 64 cells and de-duplicates with
// sort + unique.  The gfx950 form inserts each body once, into the cell of its
// AABB's min corner, and queries half of the neighbour cells (at most 14):
//   * cell side c >= every in-grid body's largest side (widened by 2^-9), so
//     two overlapping in-grid bodies have min corners less than one cell
//     apart on every axis and their home cells are neighbours;
//   * bodies larger than the cell (or with cell coordinates out of range) go
//     to an overflow list and are tested against every body, as the
//     reference's overflow path does (broadphase.cpp:162-166);
//   * home-cell keys are 3-D Morton codes of the cell coordinates wrapped to
//     k bits per axis (table of 2^3k entries, 2^3k >= 2N): spatially close
//     cells are close in memory; far cells that wrap onto the same key only
//     add AABB tests, never change the result;
//   * of two overlapping bodies with different home cells, the one whose cell
//     comes first in (z, y, x) order finds the pair (the other cell is at an
//     offset after (0, 0, 0) in that order, one of the 13 queried); in a
//     shared home cell the lower original index does; so each unordered pair
//     appears once (the 27 neighbour keys are distinct for k >= 2, so a
//     wrapped key never brings back the other half, and every body sits in
//     exactly one cell).
// Cell coordinates use floor(min / c) in double precision, a monotone map.
// Batches of worlds: every world has its own cells.  A world's keys are its
// Morton keys XOR a per-world hash (a bijection, so one world's 27 neighbour
// keys stay distinct); records of other worlds that land on the same key are
// rejected by the world test.
// Automatic cell: the largest AABB side, capped at 8x the mean side (the
// rest overflow), widened by 2^-8 (in-grid bodies must be at most 1 - 2^-9
// of the cell).
#include <stddef.h>

#include <algorithm>

#include "devmath.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace pg {
namespace {

__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t cell_key(int32_t x, int32_t y, int32_t z, uint32_t mask)
{
    return spread3(uint32_t(x) & mask) | (spread3(uint32_t(y) & mask) << 1) |
           (spread3(uint32_t(z) & mask) << 2);
}

// Per-world key offset (0 for world 0): multiplicative hash to table bits.
__device__ __forceinline__ uint32_t world_key(uint32_t world, uint32_t bits_per_axis)
{
    return world ? (world * 0x9E3779B1u) >> (32u - 3u * bits_per_axis) : 0u;
}

// Cell side for this step: explicit (> 0) or automatic from the reductions.
__device__ __forceinline__ float step_cell(const Counters* c, float requested, uint32_t n)
{
    if (requested > 0.0f)
        return requested;
    const float wmax = key_float(c->wmax_key);
    const double mean = n ? double(c->width_sum_fx) / 65536.0 / double(n) : 1.0;
    float cell = fminf(wmax, float(8.0 * mean));
    if (!(cell > 0.0f))
        cell = 1.0f;
    return cell * (1.0f + 1.0f / 256.0f);
}

// Home cell of a body; false when it must use the overflow path.
__device__ __forceinline__ bool home_cell(float x, float y, float z, double inv, int32_t cell[3])
{
    const double v[3] = {floor(double(x) * inv), floor(double(y) * inv), floor(double(z) * inv)};
    for (int a = 0; a < 3; ++a) {
        if (!(v[a] > -1073741824.0 && v[a] < 1073741824.0))
            return false;
        cell[a] = int32_t(v[a]);
    }
    return true;
}

// Key of body i (i < n): its home cell's key, or 1 << 3k for the overflow
// list; writes the sort item, the in-grid flag and, for overflow bodies,
// their list entry.  Every lane of the wave calls it.
__device__ __forceinline__ uint32_t grid_key(uint32_t n, const float4* __restrict__ aabb,
                                             Counters* counters, float cell,
                                             uint32_t bits_per_axis, uint64_t* __restrict__ items,
                                             uint32_t* __restrict__ overflow,
                                             uint32_t* __restrict__ flag, uint32_t i)
{
    bool in_grid = false;
    uint32_t key = 1u << (3 * bits_per_axis);  // sorts after every cell key
    if (i < n) {
        const Rec r = load_rec(aabb, i);
        const float limit = cell * (1.0f - 1.0f / 512.0f);
        const bool fits = (r.b[3] - r.b[0]) <= limit && (r.b[4] - r.b[1]) <= limit &&
                          (r.b[5] - r.b[2]) <= limit;
        int32_t c[3], top[3];
        const double inv = 1.0 / double(cell);
        if (fits && home_cell(r.b[0], r.b[1], r.b[2], inv, c) &&
            home_cell(r.b[3], r.b[4], r.b[5], inv, top)) {
            in_grid = true;
            key = cell_key(c[0], c[1], c[2], (1u << bits_per_axis) - 1u) ^
                  world_key(r.world, bits_per_axis);
        }
        items[i] = (uint64_t)key | ((uint64_t)i << 32);
        flag[i] = in_grid ? 1u : 0u;
    }
    // Overflow list (rare: one atomic per wave that has an overflow body).
    const bool over = i < n && !in_grid;
    const uint64_t om = __ballot(over);
    const uint32_t lane = lane_id();
    if (om) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)om) - 1u;
        uint32_t base = 0;
        if (lane == leader)
            base = atomicAdd(&counters->overflow_count, (uint32_t)__popcll(om));
        base = __shfl(base, (int)leader);
        if (over)
            overflow[base + __popcll(om & lanes_below(lane))] = i;
    }
    return key;
}

// One workgroup per radix-sort tile (tile_items keys): the keys, and the
// tile's histogram of the sort's first digit (key bits 0..7) in the sort's
// layout (hist[digit * tiles + tile]), so the sort skips its first upsweep.
__global__ __launch_bounds__(kBlock) void k_grid_keys(uint32_t n, const float4* __restrict__ aabb,
                                                      Counters* counters, float requested,
                                                      uint32_t bits_per_axis,
                                                      uint64_t* __restrict__ items,
                                                      uint32_t* __restrict__ overflow,
                                                      uint32_t* __restrict__ flag,
                                                      uint32_t tile_items,
                                                      uint32_t* __restrict__ hist)
{
    __shared__ uint32_t counts[256];
    static_assert(kBlock == 256, "one histogram bin per thread");
    const float cell = step_cell(counters, requested, n);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        counters->cell_bits = __float_as_uint(cell);
    counts[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t r = 0; r < tile_items; r += kBlock) {
        const uint32_t i = blockIdx.x * tile_items + r + threadIdx.x;
        const uint32_t key = grid_key(n, aabb, counters, cell, bits_per_axis, items, overflow,
                                      flag, i);
        if (i < n)
            atomicAdd(&counts[key & 0xFFu], 1u);
    }
    __syncthreads();
    const uint32_t tiles = (n + tile_items - 1) / tile_items;
    hist[threadIdx.x * tiles + blockIdx.x] = counts[threadIdx.x];
}

// Bodies in cells = all bodies minus the overflow ones.
__global__ __launch_bounds__(kBlock) void k_fill(uint32_t* __restrict__ p, uint64_t count,
                                                 uint32_t value)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// Cell ranges over the sorted in-grid prefix [0, entries): table[2 key] =
// first position, table[2 key + 1] = one past the last (0, 0 = empty).
// In-grid bodies sort before the overflow ones (their keys carry the top bit).
__global__ __launch_bounds__(kBlock) void k_cell_ranges(uint32_t n, const uint64_t* __restrict__ items,
                                                        Counters* counters,
                                                        uint32_t* __restrict__ table)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t entries = n - counters->overflow_count;
    if (p == 0)
        counters->grid_entries = entries;
    if (p >= entries)
        return;
    const uint32_t key = uint32_t(items[p]);
    if (p == 0 || uint32_t(items[p - 1]) != key)
        table[2 * size_t(key)] = p;
    if (p + 1 == entries || uint32_t(items[p + 1]) != key)
        table[2 * size_t(key) + 1] = p + 1;
}

// Candidates of a body: bodies whose home cell lies, on every axis, in
// [home - 1, floor(max / cell)] (a partner's min is below this body's max and
// less than one cell below its min), i.e. 2 or 3 cells per axis.
__global__ __launch_bounds__(kBlock) void k_grid_query(uint32_t n, const float4* __restrict__ s,
                                                       const uint32_t* __restrict__ table,
                                                       uint32_t bits_per_axis,
                                                       Counters* counters,
                                                       uint64_t* __restrict__ out, uint64_t cap,
                                                       uint64_t* __restrict__ tests_part)
{
    __shared__ uint64_t queue_lds[kBlock / 64][kWaveQueue];
    WaveQueue queue{queue_lds[threadIdx.x >> 6], 0};
    const uint32_t entries = n - counters->overflow_count;
    const float cell = __uint_as_float(counters->cell_bits);
    const double inv = 1.0 / double(cell);
    const uint32_t mask = (1u << bits_per_axis) - 1u;
    const uint2* range = reinterpret_cast<const uint2*>(table);
    unsigned long long tests = 0;
    // Grid-stride over sorted positions; all lanes stay in the loops together
    // so that the wave-wide append is executed convergently.
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (entries + stride - 1) / stride;
    const uint32_t block = xcd_block(blockIdx.x, gridDim.x);
    for (uint32_t round = 0; round < rounds; ++round) {
        const uint32_t p = round * stride + block * blockDim.x + threadIdx.x;
        const bool active = p < entries;
        Rec me{{0, 0, 0, 0, 0, 0}, 0, 0};
        int32_t c[3] = {0, 0, 0}, top[3] = {0, 0, 0};  // both in range for in-grid bodies
        uint32_t wk = 0;
        if (active) {
            me = load_rec(s, p);
            home_cell(me.b[0], me.b[1], me.b[2], inv, c);
            home_cell(me.b[3], me.b[4], me.b[5], inv, top);
            wk = world_key(me.world, bits_per_axis);
        }
        // Half neighbourhood: of two overlapping bodies in different cells,
        // the one whose home cell comes first in (z, y, x) order finds the
        // pair (the other's cell is at offset (dx, dy, dz) > (0, 0, 0) in
        // that order, and within its [home - 1, top] range); in the home
        // cell itself the lower index does.  So the slab dz = -1 and the
        // first four cells of the slab dz = 0 are never read.
        // One z-slab of 3 x 3 neighbour cells at a time: the range loads are
        // issued together, then each lane walks the candidates of the slab's
        // cells in one loop, so the wave waits for one table round trip and
        // max(candidates) record loads per slab instead of one round trip and
        // one loop per cell.
        for (int dz = 0; dz <= 1; ++dz) {
            uint2 r[9];
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int dx = j % 3 - 1, dy = j / 3 - 1;
                r[j] = make_uint2(0u, 0u);
                if (active && (dz > 0 || j >= 4) && c[0] + dx <= top[0] && c[1] + dy <= top[1] &&
                    c[2] + dz <= top[2])
                    r[j] = range[cell_key(c[0] + dx, c[1] + dy, c[2] + dz, mask) ^ wk];
            }
            uint32_t q = 0, qe = 0, next = 0;
            for (;;) {
                if (q >= qe && next < 9) {
                    // Next non-empty cell of the slab (static indices keep r in registers).
                    uint32_t nq = 0, nqe = 0, nn = 9;
#pragma unroll
                    for (int j = 8; j >= 0; --j)
                        if (uint32_t(j) >= next && r[j].y > r[j].x) {
                            nq = r[j].x;
                            nqe = r[j].y;
                            nn = j + 1;
                        }
                    q = nq;
                    qe = nqe;
                    next = nn;
                }
                const bool more = q < qe;
                if (__ballot(more) == 0)
                    break;
                bool hit = false;
                uint32_t b = 0;
                if (more) {
                    const Rec o = load_rec(s, q);
                    b = o.index;
                    const bool home = dz == 0 && next == 5;  // the walk is in cell j = 4
                    if ((!home || me.index < b) && me.world == o.world) {
                        ++tests;
                        hit = rec_overlap(me.b, o.b);
                    }
                    ++q;
                }
                wq_push(queue, hit, pair_value(me.index, b), out, &counters->pair_count, cap);
            }
        }
    }
    wq_flush_block(queue, out, &counters->pair_count, cap);
    block_partial(tests, tests_part);
}

// Overflow bodies against every body (broadphase.cpp:163-166): pair (o, j)
// is tested when j != o and (j is in the grid or o < j).
__global__ __launch_bounds__(kBlock) void k_grid_overflow(uint32_t n, const float4* __restrict__ aabb,
                                                          const uint32_t* __restrict__ overflow,
                                                          const uint32_t* __restrict__ flag,
                                                          Counters* counters,
                                                          uint64_t* __restrict__ out, uint64_t cap,
                                                          uint64_t* __restrict__ tests_part)
{
    __shared__ uint64_t queue_lds[kBlock / 64][kWaveQueue];
    WaveQueue queue{queue_lds[threadIdx.x >> 6], 0};
    const uint32_t count = counters->overflow_count;
    unsigned long long tests = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;
    for (uint32_t k = blockIdx.y; k < count; k += gridDim.y) {
        const uint32_t o = overflow[k];
        const Rec ro = load_rec(aabb, o);
        for (uint32_t round = 0; round < rounds; ++round) {
            const uint32_t j = round * stride + blockIdx.x * blockDim.x + threadIdx.x;
            bool hit = false;
            if (j < n && j != o && (flag[j] || o < j)) {
                const Rec rj = load_rec(aabb, j);
                if (rj.world == ro.world) {
                    ++tests;
                    hit = rec_overlap(ro.b, rj.b);
                }
            }
            wq_push(queue, hit, pair_value(o, j), out, &counters->pair_count, cap);
        }
    }
    wq_flush_block(queue, out, &counters->pair_count, cap);
    block_partial(tests, tests_part + blockIdx.y * gridDim.x);
}

}  // namespace

void launch_grid(physgpu_context* ctx, double cell_size)
{
    const uint32_t n = ctx->n;
    const uint32_t bpa = ctx->table_bits / 3;
    const uint32_t table_size = 1u << ctx->table_bits;
    Timer t;
    time_begin(ctx, "grid_keys", t);
    const uint32_t tile_items = kSortThreads * ctx->tuning.sort_items;
    hipLaunchKernelGGL(k_grid_keys, dim3(div_up(n, tile_items)), dim3(kBlock), 0, ctx->stream, n,
                       ctx->aabb, ctx->counters, float(cell_size), bpa, ctx->items_a,
                       ctx->overflow_list, ctx->grid_flag, tile_items, ctx->histogram);
    time_end(ctx, t);
    uint64_t* sorted = nullptr;
    radix_sort(ctx, ctx->items_a, ctx->items_b, n, nullptr, 0, 3 * bpa + 1, 0, 0, &sorted,
               ctx->aabb, ctx->sorted_aabb, true);
    time_begin(ctx, "grid_table_clear", t);
    hipLaunchKernelGGL(k_fill, dim3(stride_grid(2ull * table_size)), dim3(kBlock), 0, ctx->stream,
                       ctx->cell_table, 2ull * table_size, 0u);
    time_end(ctx, t);
    time_begin(ctx, "grid_cell_ranges", t);
    hipLaunchKernelGGL(k_cell_ranges, dim3(div_up(n, kBlock)), dim3(kBlock), 0, ctx->stream, n,
                       sorted, ctx->counters, ctx->cell_table);
    time_end(ctx, t);
    time_begin(ctx, "grid_query", t);
    hipLaunchKernelGGL(k_grid_query, dim3(ctx->tuning.query_groups),
                       dim3(ctx->tuning.query_block), 0, ctx->stream, n,
                       ctx->sorted_aabb, ctx->cell_table, bpa, ctx->counters, ctx->pairs_raw,
                       ctx->cap_p, ctx->partials + size_t(kSlotGridTests) * kMaxPartialBlocks);
    time_end(ctx, t);
    time_begin(ctx, "grid_overflow", t);
    // Overflow bodies are rare; a small grid keeps the idle launch cheap.
    const dim3 og(std::min<uint32_t>(stride_grid(n), 512), 8);
    hipLaunchKernelGGL(k_grid_overflow, og, dim3(kBlock), 0, ctx->stream, n, ctx->aabb,
                       ctx->overflow_list, ctx->grid_flag, ctx->counters, ctx->pairs_raw,
                       ctx->cap_p, ctx->partials + size_t(kSlotGridOverflowTests) * kMaxPartialBlocks);
    time_end(ctx, t);
    FoldList fold{};
    fold.e[0] = {kSlotGridTests, ctx->tuning.query_groups, kFoldSum64,
                 uint32_t(offsetof(Counters, grid_tests))};
    fold.e[1] = {kSlotGridOverflowTests, og.x * og.y, kFoldSum64,
                 uint32_t(offsetof(Counters, grid_overflow_tests))};
    fold.count = 2;
    launch_fold(ctx, fold);
}

}  // namespace pg
