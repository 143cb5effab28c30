// Canonical pair order (src/collision/broadphase.cpp:18-60).
//
// Every broad phase appends unordered (first, second) records, first < second,
// as first << 32 | second.  The reference orders them lexicographically with
// two stable counting passes, by `second` then by `first` (sortPairsLinear).
// Pairs are distinct, so the lexicographic order is unique and any method
// that sorts by (first, second) reproduces it.  Here one counting pass by
// `first` puts every pair in its body's segment and a segmented sort by
// `second` finishes each segment:
//   1. rank: each pair takes a slot in its `first` segment (an atomic on that
//      body's counter; counters of different bodies do not contend);
//   2. exclusive scan of the n + 1 segment counters (tile sums, one-workgroup
//      scan of the tile sums, per-tile scan);
//   3. scatter every pair to its segment;
//   4. sort each segment by `second`: a register sorting network by one thread
//      for up to 16 pairs (nearly every body), a rank sort by one wave up to
//      64, a bitonic sort in LDS by one workgroup up to 4096, and beyond that
//      one workgroup marking `second` values in an LDS bitmap window by window
//      (distinct keys: a key's position is the number of set bits before it).
// Alongside the pairs the sort writes perm: the emission index of the pair at
// each sorted position, so a narrow phase that ran on the emitted (unsorted)
// pairs can be read out in canonical order.  The pair count is read on the
// device, so nothing waits for the host; every kernel runs on the stream the
// caller names.
#include "internal.hpp"
#include "wave.hpp"

namespace pg {
namespace {

constexpr uint32_t kRegSegment = 16;
constexpr uint32_t kWaveSegment = 64;
constexpr uint32_t kLdsSegment = 4096;
constexpr uint32_t kBitmapWords = 4096;  // 2^17 keys per window
constexpr uint32_t kTile = 2048;         // scan tile: 256 threads x 8
constexpr uint32_t kRestGroups = 512;

__global__ void k_finish_pairs(Counters* c, uint64_t cap)
{
    const unsigned long long p = c->pair_count;
    c->sort_count = uint32_t(p < cap ? p : cap);
    c->mid_segments = 0;
    c->long_segments = 0;
}

__global__ __launch_bounds__(kBlock) void k_zero_u32(uint32_t* __restrict__ p, uint64_t count)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Pairs held: the emitted count clamped to the capacity (sort_count may be
// written by a kernel that runs concurrently on the main stream).
__device__ __forceinline__ uint32_t pairs_held(const Counters* c, uint64_t cap)
{
    const unsigned long long p = c->pair_count;
    return uint32_t(p < cap ? p : cap);
}

__global__ __launch_bounds__(kBlock) void k_pair_rank(const uint64_t* __restrict__ pairs,
                                                      const Counters* __restrict__ c,
                                                      uint64_t cap,
                                                      uint32_t* __restrict__ seg,
                                                      uint32_t* __restrict__ rank)
{
    const uint32_t count = pairs_held(c, cap);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < count;
         p += gridDim.x * blockDim.x)
        rank[p] = atomicAdd(&seg[uint32_t(pairs[p] >> 32)], 1u);
}

// Sum of each tile of kTile counters.
__global__ __launch_bounds__(256) void k_tile_sum(const uint32_t* __restrict__ v, uint32_t n,
                                                  uint32_t* __restrict__ tiles)
{
    __shared__ uint32_t part[4];
    const uint32_t base = blockIdx.x * kTile;
    uint32_t s = 0;
    for (uint32_t i = threadIdx.x; i < kTile; i += 256)
        if (base + i < n)
            s += v[base + i];
    for (uint32_t off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off);
    if (lane_id() == 0)
        part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        tiles[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of one value per thread over a 256-thread workgroup;
// `total` gets the sum.  Every thread must call it.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* wave_tot, uint32_t& total)
{
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= off)
            incl += t;
    }
    if (lane == 63)
        wave_tot[wave] = incl;
    __syncthreads();
    uint32_t run = incl - v;
    for (uint32_t w = 0; w < wave; ++w)
        run += wave_tot[w];
    total = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
    __syncthreads();
    return run;
}

// Up to this many tiles each workgroup of k_tile_apply sums the earlier
// tiles' totals itself (O(tiles^2) loads over the grid, cheaper than one more
// launch at these sizes); past it k_tile_offsets scans the totals first.
constexpr uint32_t kDirectTiles = 1024;

// Exclusive scan, in place, of `count` tile totals by one workgroup.
__global__ __launch_bounds__(256) void k_tile_offsets(uint32_t* __restrict__ tiles, uint32_t count)
{
    __shared__ uint32_t wave_tot[4];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < count; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t x = i < count ? tiles[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive(x, wave_tot, total);
        if (i < count)
            tiles[i] = carry + ex;
        carry += total;
    }
}

// In-place exclusive scan of tile blockIdx.x plus its offset: the sum of the
// earlier tiles' totals, reduced by the workgroup itself (direct), or read
// from the scanned totals.
__global__ __launch_bounds__(256) void k_tile_apply(uint32_t* __restrict__ v, uint32_t n,
                                                    const uint32_t* __restrict__ tiles,
                                                    bool direct)
{
    __shared__ uint32_t wave_tot[4];
    constexpr uint32_t kPer = kTile / 256;
    uint32_t all = 0;
    if (direct) {
        uint32_t before = 0;
        for (uint32_t t = threadIdx.x; t < blockIdx.x; t += 256)
            before += tiles[t];
        block_exclusive(before, wave_tot, all);
    } else {
        all = tiles[blockIdx.x];
    }
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kPer;
    uint32_t x[kPer], mine = 0;
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
        x[i] = base + i < n ? v[base + i] : 0u;
        mine += x[i];
    }
    uint32_t total;
    uint32_t run = all + block_exclusive(mine, wave_tot, total);
#pragma unroll
    for (uint32_t i = 0; i < kPer; ++i) {
        if (base + i < n)
            v[base + i] = run;
        run += x[i];
    }
}

__global__ __launch_bounds__(kBlock) void k_pair_scatter(const uint64_t* __restrict__ pairs,
                                                         const Counters* __restrict__ c,
                                                         uint64_t cap,
                                                         const uint32_t* __restrict__ seg,
                                                         const uint32_t* __restrict__ rank,
                                                         uint64_t* __restrict__ out,
                                                         uint32_t* __restrict__ perm)
{
    const uint32_t count = pairs_held(c, cap);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < count;
         p += gridDim.x * blockDim.x) {
        const uint64_t v = pairs[p];
        const uint32_t pos = seg[uint32_t(v >> 32)] + rank[p];
        out[pos] = v;
        perm[pos] = p;
    }
}

// Append `value` to list[counter++] for lanes with pred, one atomic per wave.
// All lanes must call it.
__device__ __forceinline__ void wave_push(bool pred, uint32_t value, uint32_t* __restrict__ list,
                                          uint32_t* counter)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0)
        return;
    const uint32_t lane = lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    if (pred)
        list[base + (uint32_t)__popcll(mask & lanes_below(lane))] = value;
}

__device__ __forceinline__ void compare_swap(uint64_t& a, uint64_t& b)
{
    const bool swap = b < a;
    const uint64_t lo = swap ? b : a, hi = swap ? a : b;
    a = lo;
    b = hi;
}

// Bitonic sorting network over K registers (K a power of two), ascending.
// Entries are second << 32 | emission index: seconds are distinct within a
// segment, so this orders by second and carries the index along.
template <uint32_t K>
__device__ __forceinline__ void reg_sort(uint64_t (&k)[K])
{
#pragma unroll
    for (uint32_t size = 2; size <= K; size <<= 1)
#pragma unroll
        for (uint32_t j = size >> 1; j > 0; j >>= 1)
#pragma unroll
            for (uint32_t x = 0; x < K; ++x) {
                const uint32_t y = x ^ j;
                if (y > x) {
                    if ((x & size) == 0)
                        compare_swap(k[x], k[y]);
                    else
                        compare_swap(k[y], k[x]);
                }
            }
}

// Sort out[b, b + len) (len <= K, all with the same `first`) by `second`,
// perm alongside.
template <uint32_t K>
__device__ __forceinline__ void sort_in_registers(uint64_t* __restrict__ out,
                                                  uint32_t* __restrict__ perm, uint32_t b,
                                                  uint32_t len, uint64_t hi)
{
    uint64_t k[K];
#pragma unroll
    for (uint32_t x = 0; x < K; ++x)
        k[x] = x < len ? (out[b + x] << 32) | perm[b + x] : ~0ull;
    reg_sort<K>(k);
#pragma unroll
    for (uint32_t x = 0; x < K; ++x)
        if (x < len) {
            out[b + x] = hi | (k[x] >> 32);
            perm[b + x] = uint32_t(k[x]);
        }
}

// One thread per body: segments of up to kRegSegment pairs are sorted in
// registers; longer ones are listed for k_sort_rest.
__global__ __launch_bounds__(kBlock) void k_sort_short(uint32_t n, const uint32_t* __restrict__ seg,
                                                       uint64_t* __restrict__ out,
                                                       uint32_t* __restrict__ perm, Counters* c,
                                                       uint32_t* __restrict__ mid_list,
                                                       uint32_t* __restrict__ long_list)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        uint32_t b = 0, len = 0;
        if (i < n) {
            b = seg[i];
            len = seg[i + 1] - b;
        }
        const uint64_t hi = uint64_t(i) << 32;
        const bool reg = len > 1 && len <= kRegSegment;
        if (__ballot(reg && len > 4)) {
            if (reg)
                sort_in_registers<kRegSegment>(out, perm, b, len, hi);
        } else if (reg) {
            sort_in_registers<4>(out, perm, b, len, hi);
        }
        wave_push(len > kRegSegment && len <= kWaveSegment, i, mid_list, &c->mid_segments);
        wave_push(len > kWaveSegment, i, long_list, &c->long_segments);
    }
}

// Segments listed by k_sort_short.  Waves take the mid segments (rank sort by
// one wave: lane x holds key x, its rank is the number of smaller keys), then
// workgroups take the long ones: a bitonic sort of second << 32 | index in
// LDS, or beyond kLdsSegment bitmap windows (a key's rank is the number of
// set bits below it: per-word prefix counts plus a popcount), written through
// the scratch arrays at the segment's own range.
__global__ __launch_bounds__(256) void k_sort_rest(const uint32_t* __restrict__ seg,
                                                   const Counters* __restrict__ c,
                                                   const uint32_t* __restrict__ mid_list,
                                                   const uint32_t* __restrict__ long_list,
                                                   uint64_t* __restrict__ out,
                                                   uint32_t* __restrict__ perm,
                                                   uint64_t* __restrict__ scratch,
                                                   uint32_t* __restrict__ scratch_perm)
{
    __shared__ uint64_t lds[kLdsSegment];  // bitonic keys, or bitmap words + their prefixes
    __shared__ uint32_t wave_tot[4];
    __shared__ uint32_t bounds[8];
    uint32_t* bits = reinterpret_cast<uint32_t*>(lds);
    uint32_t* prefix = bits + kBitmapWords;

    const uint32_t lane = lane_id();
    const uint32_t mids = c->mid_segments;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); s < mids; s += waves) {
        const uint32_t i = mid_list[s];
        const uint32_t b = seg[i], len = seg[i + 1] - b;
        const uint32_t key = lane < len ? uint32_t(out[b + lane]) : 0xFFFFFFFFu;
        const uint32_t idx = lane < len ? perm[b + lane] : 0u;
        uint32_t r = 0;
#pragma unroll
        for (uint32_t j = 0; j < 64; ++j)
            r += lane_value(key, j) < key ? 1u : 0u;
        if (lane < len) {
            out[b + r] = (uint64_t(i) << 32) | key;
            perm[b + r] = idx;
        }
    }

    const uint32_t longs = c->long_segments;
    for (uint32_t s = blockIdx.x; s < longs; s += gridDim.x) {
        const uint32_t i = long_list[s];
        const uint32_t b = seg[i], len = seg[i + 1] - b;
        const uint64_t hi = uint64_t(i) << 32;
        if (len <= kLdsSegment) {
            uint32_t size = 1;
            while (size < len)
                size <<= 1;
            for (uint32_t k = threadIdx.x; k < size; k += blockDim.x)
                lds[k] = k < len ? (out[b + k] << 32) | perm[b + k] : ~0ull;
            __syncthreads();
            for (uint32_t k = 2; k <= size; k <<= 1)
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t x = threadIdx.x; x < size; x += blockDim.x) {
                        const uint32_t y = x ^ j;
                        if (y > x) {
                            const uint64_t kx = lds[x], ky = lds[y];
                            if ((kx > ky) == ((x & k) == 0)) {
                                lds[x] = ky;
                                lds[y] = kx;
                            }
                        }
                    }
                    __syncthreads();
                }
            for (uint32_t k = threadIdx.x; k < len; k += blockDim.x) {
                out[b + k] = hi | (lds[k] >> 32);
                perm[b + k] = uint32_t(lds[k]);
            }
            __syncthreads();
            continue;
        }
        // Key range of the segment.
        uint32_t mn = 0xFFFFFFFFu, mx = 0;
        for (uint32_t k = threadIdx.x; k < len; k += blockDim.x) {
            const uint32_t key = uint32_t(out[b + k]);
            mn = min(mn, key);
            mx = max(mx, key);
        }
        for (uint32_t off = 32; off > 0; off >>= 1) {
            mn = min(mn, (uint32_t)__shfl_xor(mn, off));
            mx = max(mx, (uint32_t)__shfl_xor(mx, off));
        }
        if (lane == 0) {
            bounds[threadIdx.x >> 6] = mn;
            bounds[4 + (threadIdx.x >> 6)] = mx;
        }
        __syncthreads();
        mn = min(min(bounds[0], bounds[1]), min(bounds[2], bounds[3]));
        mx = max(max(bounds[4], bounds[5]), max(bounds[6], bounds[7]));
        __syncthreads();
        constexpr uint32_t kWindow = kBitmapWords * 32;
        constexpr uint32_t kPer = kBitmapWords / 256;
        uint32_t placed = 0;
        for (uint64_t w0 = mn; w0 <= mx; w0 += kWindow) {
            for (uint32_t k = threadIdx.x; k < kBitmapWords; k += blockDim.x)
                bits[k] = 0;
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < len; k += blockDim.x) {
                const uint64_t rel = uint64_t(uint32_t(out[b + k])) - w0;
                if (rel < kWindow)
                    atomicOr(&bits[rel >> 5], 1u << (rel & 31u));
            }
            __syncthreads();
            // Exclusive prefix of the set bits per word.
            uint32_t cnt = 0;
#pragma unroll
            for (uint32_t q = 0; q < kPer; ++q)
                cnt += __popc(bits[threadIdx.x * kPer + q]);
            uint32_t total;
            uint32_t run = placed + block_exclusive(cnt, wave_tot, total);
#pragma unroll
            for (uint32_t q = 0; q < kPer; ++q) {
                prefix[threadIdx.x * kPer + q] = run;
                run += __popc(bits[threadIdx.x * kPer + q]);
            }
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < len; k += blockDim.x) {
                const uint32_t key = uint32_t(out[b + k]);
                const uint64_t rel = uint64_t(key) - w0;
                if (rel < kWindow) {
                    const uint32_t w = uint32_t(rel >> 5), bit = uint32_t(rel & 31u);
                    const uint32_t pos = prefix[w] + __popc(bits[w] & ((1u << bit) - 1u));
                    scratch[b + pos] = hi | key;
                    scratch_perm[b + pos] = perm[b + k];
                }
            }
            placed += total;
            __syncthreads();
        }
        for (uint32_t k = threadIdx.x; k < len; k += blockDim.x) {
            out[b + k] = scratch[b + k];
            perm[b + k] = scratch_perm[b + k];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_swap_halves(uint64_t* __restrict__ p, uint64_t count)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = p[i];
        p[i] = (v >> 32) | (v << 32);
    }
}

__global__ void k_set_pair_count(Counters* c, uint64_t count)
{
    c->pair_count = count;
    c->sort_count = uint32_t(count);
}

}  // namespace

namespace {

// In-place exclusive scan of n u32 values on `stream` (tile sums in seg_tiles).
void scan_u32(physgpu_context* ctx, uint32_t* v, uint32_t n, hipStream_t stream)
{
    const uint32_t tiles = div_up(n, kTile);
    hipLaunchKernelGGL(k_tile_sum, dim3(tiles), dim3(256), 0, stream, v, n, ctx->seg_tiles);
    const bool direct = tiles <= kDirectTiles;
    if (!direct)
        hipLaunchKernelGGL(k_tile_offsets, dim3(1), dim3(256), 0, stream, ctx->seg_tiles, tiles);
    hipLaunchKernelGGL(k_tile_apply, dim3(tiles), dim3(256), 0, stream, v, n, ctx->seg_tiles,
                       direct);
}

}  // namespace

void launch_finish_pairs(physgpu_context* ctx)
{
    hipLaunchKernelGGL(k_finish_pairs, dim3(1), dim3(1), 0, ctx->stream, ctx->counters,
                       ctx->cap_p);
}

void launch_sort_pairs(physgpu_context* ctx, hipStream_t stream)
{
    Timer t;
    const uint32_t n = ctx->n;
    const uint32_t segs = n + 1;
    const uint32_t pgrid = stride_grid(ctx->cap_p);
    time_begin(ctx, "pair_rank", t, stream);
    hipLaunchKernelGGL(k_zero_u32, dim3(stride_grid(segs)), dim3(kBlock), 0, stream,
                       ctx->seg_off, uint64_t(segs));
    hipLaunchKernelGGL(k_pair_rank, dim3(pgrid), dim3(kBlock), 0, stream, ctx->pairs_raw,
                       ctx->counters, ctx->cap_p, ctx->seg_off, ctx->pair_rank);
    time_end(ctx, t);
    time_begin(ctx, "pair_scan", t, stream);
    scan_u32(ctx, ctx->seg_off, segs, stream);
    time_end(ctx, t);
    time_begin(ctx, "pair_scatter", t, stream);
    hipLaunchKernelGGL(k_pair_scatter, dim3(pgrid), dim3(kBlock), 0, stream, ctx->pairs_raw,
                       ctx->counters, ctx->cap_p, ctx->seg_off, ctx->pair_rank, ctx->pairs_tmp,
                       ctx->pair_perm);
    time_end(ctx, t);
    time_begin(ctx, "pair_segment_sort", t, stream);
    hipLaunchKernelGGL(k_sort_short, dim3(stride_grid(n)), dim3(kBlock), 0, stream, n,
                       ctx->seg_off, ctx->pairs_tmp, ctx->pair_perm, ctx->counters,
                       ctx->seg_lists, ctx->seg_lists + ctx->cap_n);
    hipLaunchKernelGGL(k_sort_rest, dim3(kRestGroups), dim3(256), 0, stream, ctx->seg_off,
                       ctx->counters, ctx->seg_lists, ctx->seg_lists + ctx->cap_n, ctx->pairs_tmp,
                       ctx->pair_perm, ctx->sort_scratch, ctx->sort_scratch_perm);
    time_end(ctx, t);
    ctx->pairs = ctx->pairs_tmp;
}

void launch_canonical_pairs(physgpu_context* ctx)
{
    launch_finish_pairs(ctx);
    launch_sort_pairs(ctx, ctx->stream);
}

// physgpu_pair records (first, second) already copied into pairs_raw as raw
// 8-byte words become first << 32 | second in place.
void launch_pairs_in(physgpu_context* ctx, const physgpu_pair* /*dev_in*/, uint64_t count)
{
    if (count)
        hipLaunchKernelGGL(k_swap_halves, dim3(stride_grid(count)), dim3(kBlock), 0, ctx->stream,
                           ctx->pairs_raw, count);
    hipLaunchKernelGGL(k_set_pair_count, dim3(1), dim3(1), 0, ctx->stream, ctx->counters, count);
    ctx->pairs = ctx->pairs_raw;
}

// Canonical pairs to physgpu_pair layout in `dev_out` (device memory).
void launch_pairs_out(physgpu_context* ctx, physgpu_pair* dev_out, uint64_t count)
{
    if (!count)
        return;
    hipMemcpyAsync(dev_out, ctx->pairs, count * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                   ctx->stream);
    hipLaunchKernelGGL(k_swap_halves, dim3(stride_grid(count)), dim3(kBlock), 0, ctx->stream,
                       reinterpret_cast<uint64_t*>(dev_out), count);
}

}  // namespace pg
