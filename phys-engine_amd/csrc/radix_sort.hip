// Stable LSD radix sort of u64 items by 8-bit digits, LDS-staged, for gfx950.
//
// Used for the SAP records (key = orderable min.x in the low word, original
// index in the high word, so equal keys keep index order exactly like the
// reference comparator, broadphase.cpp:213-217), the grid's home-cell keys,
// the LBVH Morton codes and the canonical pair order (second, then first;
// the two stable counting passes of sortPairsLinear, broadphase.cpp:27-60).
//
// One pass = three kernels:
//   upsweep    per-tile digit histogram (LDS atomics), digit-major layout
//   scan       one workgroup per digit scans that digit's tile counts
//   downsweep  per-wave stable ranking with 8 ballots per item (64-wide
//              wavefronts: peers = lanes whose digit matches on all 8 bits),
//              running per-wave digit counters in LDS, tile reorder in LDS,
//              then coalesced scatter of digit runs.
// The item count may live on the device (pairs emitted by the broad phase);
// every kernel reads it and idle tiles exit, so no host round trip is needed.
#include "internal.hpp"

namespace pg {
namespace {

__device__ __forceinline__ uint32_t digit_of(uint64_t item, uint32_t shift)
{
    return uint32_t(item >> shift) & 0xFFu;
}

__device__ __forceinline__ uint32_t item_count(uint32_t n, const uint32_t* dev_count)
{
    return dev_count ? min(*dev_count, n) : n;
}

template <uint32_t kItems>
__global__ __launch_bounds__(kSortThreads) void k_upsweep(const uint64_t* __restrict__ items,
                                                          uint32_t n, const uint32_t* dev_count,
                                                          uint32_t shift, uint64_t drop,
                                                          uint32_t* __restrict__ hist)
{
    constexpr uint32_t kSortTile = kSortThreads * kItems;
    __shared__ uint32_t counts[256];
    const uint32_t count = item_count(n, dev_count);
    const uint32_t tiles = (count + kSortTile - 1) / kSortTile;
    const uint32_t tile = blockIdx.x;
    if (tile >= tiles)
        return;
    counts[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = tile * kSortTile;
#pragma unroll
    for (uint32_t r = 0; r < kItems; ++r) {
        uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < count) {
            const uint64_t it = items[i];
            if (!(it & drop))
                atomicAdd(&counts[digit_of(it, shift)], 1u);
        }
    }
    __syncthreads();
    hist[threadIdx.x * tiles + tile] = counts[threadIdx.x];
}

// Exclusive scan of 256 values held one per thread of a 256-thread
// workgroup (wave scans with shuffles, then the four wave totals).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t wave_sums[kSortThreads / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= off)
            incl += t;
    }
    __syncthreads();
    if (lane == 63)
        wave_sums[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < kSortThreads / 64; ++w) {
        if (w < wave)
            before += wave_sums[w];
        all += wave_sums[w];
    }
    *total = all;
    return before + incl - v;
}

// One workgroup per digit: exclusive scan of that digit's row of tile counts
// (hist[d * tiles + tile]) in place, and the digit's total.
template <uint32_t kItems>
__global__ __launch_bounds__(kSortThreads) void k_scan_rows(uint32_t* __restrict__ hist,
                                                            uint32_t n, const uint32_t* dev_count,
                                                            uint32_t* __restrict__ digit_total)
{
    constexpr uint32_t kSortTile = kSortThreads * kItems;
    const uint32_t count = item_count(n, dev_count);
    const uint32_t tiles = (count + kSortTile - 1) / kSortTile;
    uint32_t* row = hist + size_t(blockIdx.x) * tiles;
    uint32_t run = 0;
    for (uint32_t base = 0; base < tiles; base += kSortThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < tiles ? row[i] : 0u;
        uint32_t chunk;
        const uint32_t ex = block_exclusive_scan(v, &chunk);
        if (i < tiles)
            row[i] = run + ex;
        run += chunk;
    }
    if (threadIdx.x == 0)
        digit_total[blockIdx.x] = run;
}

constexpr uint32_t kWaves = kSortThreads / 64;

template <uint32_t kItems>
__global__ __launch_bounds__(kSortThreads) void k_downsweep(const uint64_t* __restrict__ in,
                                                            uint64_t* __restrict__ out,
                                                            uint32_t n,
                                                            const uint32_t* dev_count,
                                                            uint32_t shift, uint64_t drop,
                                                            const uint32_t* __restrict__ hist,
                                                            const uint32_t* __restrict__ digit_total,
                                                            const float4* __restrict__ gsrc,
                                                            float4* __restrict__ gdst,
                                                            uint32_t* __restrict__ kept_out)
{
    constexpr uint32_t kSortTile = kSortThreads * kItems;
    constexpr uint32_t kWaveItems = kSortTile / kWaves;  // contiguous sub-tile per wave
    __shared__ uint64_t tile_items[kSortTile];
    __shared__ uint32_t wave_count[kWaves][256];
    __shared__ uint32_t tile_start[256];
    __shared__ uint32_t global_start[256];
    __shared__ uint32_t kept;

    const uint32_t count = item_count(n, dev_count);
    const uint32_t tiles = (count + kSortTile - 1) / kSortTile;
    const uint32_t tile = blockIdx.x;
    if (tile >= tiles)
        return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t base = tile * kSortTile;
    const uint32_t valid = min(kSortTile, count - base);

    for (uint32_t w = 0; w < kWaves; ++w)
        wave_count[w][threadIdx.x] = 0;
    {
        // Global start of this tile's digit run: all smaller digits, then
        // this digit in earlier tiles.
        uint32_t all;
        const uint32_t digit_base = block_exclusive_scan(digit_total[threadIdx.x], &all);
        global_start[threadIdx.x] = digit_base + hist[threadIdx.x * tiles + tile];
        if (kept_out && tile == 0 && threadIdx.x == 0)
            *kept_out = all;  // items of the pass that were not dropped
    }
    __syncthreads();

    // Each wave ranks its contiguous sub-tile, 64 items at a time, in order.
    uint64_t item[kItems];
    uint32_t rank[kItems];
    uint32_t dig[kItems];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
#pragma unroll
    for (uint32_t r = 0; r < kItems; ++r) {
        const uint32_t local = wave * kWaveItems + r * 64u + lane;
        item[r] = local < valid ? in[base + local] : 0ull;
        const bool ok = local < valid && !(item[r] & drop);
        const uint32_t d = digit_of(item[r], shift);
        dig[r] = d;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        uint32_t before = 0;
        if (ok)
            before = wave_count[wave][d];
        const uint32_t below = (uint32_t)__popcll(peers & lt_mask);
        rank[r] = before + below;
        // The lowest peer publishes the new running count for this digit.
        if (ok && below == 0)
            wave_count[wave][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // Per digit: exclusive prefix over waves, then over digits.
    {
        const uint32_t d = threadIdx.x;
        uint32_t run = 0;
        for (uint32_t w = 0; w < kWaves; ++w) {
            uint32_t c = wave_count[w][d];
            wave_count[w][d] = run;
            run += c;
        }
        tile_start[d] = run;  // digit total for now
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // 256-entry exclusive scan by one wave (4 entries per lane).
        uint32_t v[4];
        uint32_t s = 0;
        for (int k = 0; k < 4; ++k) {
            v[k] = tile_start[lane * 4 + k];
            s += v[k];
        }
        uint32_t incl = s;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            uint32_t t = __shfl_up(incl, off);
            if (lane >= off)
                incl += t;
        }
        uint32_t run = incl - s;
        for (int k = 0; k < 4; ++k) {
            tile_start[lane * 4 + k] = run;
            run += v[k];
        }
        if (lane == 63)
            kept = run;  // items of the tile that are not dropped
    }
    __syncthreads();

#pragma unroll
    for (uint32_t r = 0; r < kItems; ++r) {
        const uint32_t local = wave * kWaveItems + r * 64u + lane;
        if (local < valid && !(item[r] & drop)) {
            const uint32_t d = dig[r];
            tile_items[tile_start[d] + wave_count[wave][d] + rank[r]] = item[r];
        }
    }
    __syncthreads();

    // Coalesced scatter: consecutive tile slots of one digit go to
    // consecutive global slots.
#pragma unroll
    for (uint32_t r = 0; r < kItems; ++r) {
        const uint32_t s = r * kSortThreads + threadIdx.x;
        if (s < kept) {
            const uint64_t it = tile_items[s];
            const uint32_t d = digit_of(it, shift);
            const uint32_t pos = global_start[d] + (s - tile_start[d]);
            out[pos] = it;
            if (gdst) {
                // Final pass: the record of the item's index (high word) too.
                const size_t idx = size_t(it >> 32);
                gdst[2 * size_t(pos)] = gsrc[2 * idx];
                gdst[2 * size_t(pos) + 1] = gsrc[2 * idx + 1];
            }
        }
    }
}

template <uint32_t kItems>
void radix_sort_tiled(physgpu_context* ctx, uint64_t* a, uint64_t* b, uint32_t n,
                      const uint32_t* device_count, uint32_t shift0, uint32_t bits0,
                      uint32_t shift1, uint32_t bits1, uint64_t** result,
                      const float4* gather_src, float4* gather_dst, bool first_hist_done,
                      uint64_t drop, uint32_t* kept)
{
    const uint32_t tiles = div_up(n, kSortThreads * kItems);
    const uint32_t passes = (bits0 + 7) / 8 + (bits1 + 7) / 8;
    uint32_t done = 0;
    uint64_t* src = a;
    uint64_t* dst = b;
    auto pass = [&](uint32_t shift) {
        const bool first = done == 0;
        const bool last = ++done == passes;
        const uint64_t pass_drop = first ? drop : 0ull;
        // After a pass that dropped items, the rest count only the kept ones.
        const uint32_t* count_in = first || !kept ? device_count : kept;
        Timer t;
        if (!(first && first_hist_done)) {
            time_begin(ctx, "radix_upsweep", t);
            hipLaunchKernelGGL(k_upsweep<kItems>, dim3(tiles), dim3(kSortThreads), 0,
                               ctx->stream, src, n, count_in, shift, pass_drop,
                               ctx->histogram);
            time_end(ctx, t);
        }
        time_begin(ctx, "radix_scan", t);
        hipLaunchKernelGGL(k_scan_rows<kItems>, dim3(256), dim3(kSortThreads), 0, ctx->stream,
                           ctx->histogram, n, count_in, ctx->digit_total);
        time_end(ctx, t);
        time_begin(ctx, "radix_downsweep", t);
        hipLaunchKernelGGL(k_downsweep<kItems>, dim3(tiles), dim3(kSortThreads), 0, ctx->stream,
                           src, dst, n, count_in, shift, pass_drop, ctx->histogram,
                           ctx->digit_total,
                           last ? gather_src : nullptr, last ? gather_dst : nullptr,
                           first ? kept : nullptr);
        time_end(ctx, t);
        uint64_t* tmp = src;
        src = dst;
        dst = tmp;
    };
    if (n > 0) {
        for (uint32_t s = 0; s < bits0; s += 8)
            pass(shift0 + s);
        for (uint32_t s = 0; s < bits1; s += 8)
            pass(shift1 + s);
    }
    *result = src;
}

}  // namespace

void radix_sort(physgpu_context* ctx, uint64_t* a, uint64_t* b, uint32_t n,
                const uint32_t* device_count, uint32_t shift0, uint32_t bits0, uint32_t shift1,
                uint32_t bits1, uint64_t** result, const float4* gather_src, float4* gather_dst,
                bool first_hist_done, uint64_t drop, uint32_t* kept)
{
    // Tile = 256 threads x sort_items (4, 8 or 16); the histogram is sized for
    // the smallest tile.
    switch (ctx->tuning.sort_items) {
    case 4:
        radix_sort_tiled<4>(ctx, a, b, n, device_count, shift0, bits0, shift1, bits1, result,
                                gather_src, gather_dst, first_hist_done, drop, kept);
        break;
    case 16:
        radix_sort_tiled<16>(ctx, a, b, n, device_count, shift0, bits0, shift1, bits1, result,
                                gather_src, gather_dst, first_hist_done, drop, kept);
        break;
    default:
        radix_sort_tiled<8>(ctx, a, b, n, device_count, shift0, bits0, shift1, bits1, result,
                                gather_src, gather_dst, first_hist_done, drop, kept);
        break;
    }
}

}  // namespace pg
