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
//   scan       one workgroup turns the histogram into global offsets
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

__global__ __launch_bounds__(kSortThreads) void k_upsweep(const uint64_t* __restrict__ items,
                                                          uint32_t n, const uint32_t* dev_count,
                                                          uint32_t shift,
                                                          uint32_t* __restrict__ hist)
{
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
    for (uint32_t r = 0; r < kSortItems; ++r) {
        uint32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < count)
            atomicAdd(&counts[digit_of(items[i], shift)], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + tile] = counts[threadIdx.x];
}

// Exclusive scan of the digit-major histogram (256 x tiles entries).
__global__ __launch_bounds__(1024) void k_scan(uint32_t* __restrict__ hist, uint32_t n,
                                               const uint32_t* dev_count)
{
    __shared__ uint32_t sums[1024];
    const uint32_t count = item_count(n, dev_count);
    const uint32_t tiles = (count + kSortTile - 1) / kSortTile;
    const uint32_t entries = tiles * 256;
    const uint32_t per = (entries + 1023) / 1024;
    const uint32_t begin = min(threadIdx.x * per, entries);
    const uint32_t end = min(begin + per, entries);
    uint32_t s = 0;
    for (uint32_t i = begin; i < end; ++i)
        s += hist[i];
    sums[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partial sums.
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? sums[threadIdx.x - off] : 0u;
        __syncthreads();
        sums[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? sums[threadIdx.x - 1] : 0u;
    for (uint32_t i = begin; i < end; ++i) {
        uint32_t c = hist[i];
        hist[i] = run;
        run += c;
    }
}

constexpr uint32_t kWaves = kSortThreads / 64;
constexpr uint32_t kWaveItems = kSortTile / kWaves;  // contiguous sub-tile per wave

__global__ __launch_bounds__(kSortThreads) void k_downsweep(const uint64_t* __restrict__ in,
                                                            uint64_t* __restrict__ out,
                                                            uint32_t n,
                                                            const uint32_t* dev_count,
                                                            uint32_t shift,
                                                            const uint32_t* __restrict__ hist)
{
    __shared__ uint64_t tile_items[kSortTile];
    __shared__ uint32_t wave_count[kWaves][256];
    __shared__ uint32_t tile_start[256];
    __shared__ uint32_t global_start[256];

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
    global_start[threadIdx.x] = hist[threadIdx.x * tiles + tile];
    __syncthreads();

    // Each wave ranks its contiguous sub-tile, 64 items at a time, in order.
    uint64_t item[kSortItems];
    uint32_t rank[kSortItems];
    uint32_t dig[kSortItems];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; ++r) {
        const uint32_t local = wave * kWaveItems + r * 64u + lane;
        const bool ok = local < valid;
        item[r] = ok ? in[base + local] : 0ull;
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
    }
    __syncthreads();

#pragma unroll
    for (uint32_t r = 0; r < kSortItems; ++r) {
        const uint32_t local = wave * kWaveItems + r * 64u + lane;
        if (local < valid) {
            const uint32_t d = dig[r];
            tile_items[tile_start[d] + wave_count[wave][d] + rank[r]] = item[r];
        }
    }
    __syncthreads();

    // Coalesced scatter: consecutive tile slots of one digit go to
    // consecutive global slots.
#pragma unroll
    for (uint32_t r = 0; r < kSortItems; ++r) {
        const uint32_t s = r * kSortThreads + threadIdx.x;
        if (s < valid) {
            const uint64_t it = tile_items[s];
            const uint32_t d = digit_of(it, shift);
            out[global_start[d] + (s - tile_start[d])] = it;
        }
    }
}

}  // namespace

void radix_sort(physgpu_context* ctx, uint64_t* a, uint64_t* b, uint32_t n,
                const uint32_t* device_count, uint32_t shift0, uint32_t bits0, uint32_t shift1,
                uint32_t bits1, uint64_t** result)
{
    const uint32_t tiles = div_up(n, kSortTile);
    uint64_t* src = a;
    uint64_t* dst = b;
    auto pass = [&](uint32_t shift) {
        Timer t;
        time_begin(ctx, "radix_upsweep", t);
        hipLaunchKernelGGL(k_upsweep, dim3(tiles), dim3(kSortThreads), 0, ctx->stream, src, n,
                           device_count, shift, ctx->histogram);
        time_end(ctx, t);
        time_begin(ctx, "radix_scan", t);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, ctx->stream, ctx->histogram, n,
                           device_count);
        time_end(ctx, t);
        time_begin(ctx, "radix_downsweep", t);
        hipLaunchKernelGGL(k_downsweep, dim3(tiles), dim3(kSortThreads), 0, ctx->stream, src,
                           dst, n, device_count, shift, ctx->histogram);
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

}  // namespace pg
