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
//     k bits per axis (table of 2^3k entries, 2^3k >= 2N, k >= 2): spatially
//     close cells are close in memory; far cells that wrap onto the same key
//     only add AABB tests, never change the result;
//   * of two overlapping bodies with different home cells, the one whose cell
//     comes first in (z, y, x) order finds the pair (the other cell is at an
//     offset after (0, 0, 0) in that order, one of the 13 queried); in a
//     shared home cell the lower original index does; so each unordered pair
//     appears once (the 27 neighbour keys are distinct for k >= 2: three
//     consecutive coordinates differ in their low two bits, so a wrapped key
//     never brings back the other half, and every body sits in exactly one
//     cell);
//   * the records are put in key order by buckets: one stable radix pass
//     (two past 2^21 table entries) over the key bits [low, 3k) groups the
//     keys by bucket, then one workgroup per bucket counts its 2^low cells
//     (low <= 13) in LDS, writes their table entries (start, end) -- every
//     entry of the table, so no clearing pass -- and places each record.
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
#include "gridkey.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace pg {
namespace {

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

// Key of body i (i < n): its home cell's key, or kOverflowKey for the
// overflow list; writes the sort item and, for overflow bodies, their list
// entry.  Every lane of the wave calls it.
__device__ __forceinline__ uint32_t grid_key(uint32_t n, const float4* __restrict__ aabb,
                                             Counters* counters, float cell,
                                             uint32_t bits_per_axis, uint32_t table_bits,
                                             uint64_t* __restrict__ items,
                                             uint32_t* __restrict__ overflow,
                                             uint32_t* __restrict__ over_ctr, uint32_t i)
{
    uint32_t key = kOverflowKey;
    if (i < n) {
        const Rec r = load_rec(aabb, i);
        key = grid_cell_key(r.b, rec_world(r.world), cell, 1.0 / double(cell), bits_per_axis,
                            table_bits);
        items[i] = (uint64_t)key | ((uint64_t)i << 32);
    }
    // Overflow list (rare: one atomic per wave that has an overflow body).
    const bool over = i < n && key == kOverflowKey;
    const uint64_t om = __ballot(over);
    const uint32_t lane = lane_id();
    if (om) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)om) - 1u;
        uint32_t base = 0;
        if (lane == leader)
            base = atomicAdd(over_ctr, (uint32_t)__popcll(om));
        base = __shfl(base, (int)leader);
        if (over)
            overflow[base + __popcll(om & lanes_below(lane))] = i;
    }
    return key;
}

// One workgroup per radix-sort tile (tile_items keys): the keys, and the
// tile's histogram of the bucket pass's first digit (key bits low .. low+7,
// in-grid bodies only) in the sort's layout (hist[digit * tiles + tile]), so
// the sort skips its first upsweep.
__global__ __launch_bounds__(kBlock) void k_grid_keys(uint32_t n, const float4* __restrict__ aabb,
                                                      Counters* counters, float requested,
                                                      uint32_t bits_per_axis, uint32_t table_bits,
                                                      uint32_t low,
                                                      uint64_t* __restrict__ items,
                                                      uint32_t* __restrict__ overflow,
                                                      uint32_t* __restrict__ over_ctr,
                                                      uint32_t tile_items,
                                                      uint32_t* __restrict__ hist,
                                                      uint32_t* __restrict__ bounds,
                                                      uint32_t bound_words)
{
    // Bucket ranges of a two-pass sort start empty.
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < bound_words;
         i += gridDim.x * blockDim.x)
        bounds[i] = 0;
    __shared__ uint32_t counts[256];
    static_assert(kBlock == 256, "one histogram bin per thread");
    const float cell = step_cell(counters, requested, n);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        counters->cell_bits = __float_as_uint(cell);
    counts[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t r = 0; r < tile_items; r += kBlock) {
        const uint32_t i = blockIdx.x * tile_items + r + threadIdx.x;
        const uint32_t key = grid_key(n, aabb, counters, cell, bits_per_axis, table_bits, items,
                                      overflow, over_ctr, i);
        if (i < n && key != kOverflowKey)
            atomicAdd(&counts[(key >> low) & 0xFFu], 1u);
    }
    __syncthreads();
    const uint32_t tiles = (n + tile_items - 1) / tile_items;
    hist[threadIdx.x * tiles + blockIdx.x] = counts[threadIdx.x];
}

// Bucket kernel: workgroups of 1024 threads, 4 items per thread in flight.
constexpr uint32_t kBucketThreads = 1024;
constexpr uint32_t kBucketItems = 4;

// Sum of v over the workgroup (every thread calls it; v from each thread).
__device__ __forceinline__ uint32_t bucket_block_sum(uint32_t v, uint32_t* lds_wave)
{
    for (uint32_t off = 32; off > 0; off >>= 1)
        v += (uint32_t)__shfl_xor((int)v, off);
    const uint32_t wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63u) == 0)
        lds_wave[wave] = v;
    __syncthreads();
    uint32_t all = 0;
    for (uint32_t w = 0; w < kBucketThreads / 64; ++w)
        all += lds_wave[w];
    return all;
}

// Exclusive prefix of v over the workgroup's threads in thread order.
__device__ __forceinline__ uint32_t bucket_block_scan(uint32_t v, uint32_t* lds_wave)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= off)
            incl += t;
    }
    __syncthreads();
    if (lane == 63)
        lds_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w)
        before += lds_wave[w];
    return before + incl - v;
}

// Bucket ranges after two radix passes (buckets of more than 8 bits): the
// items are sorted by bucket, so bucket b occupies [first p with bucket b,
// one past the last); each position that starts or ends a run writes it.
// Buckets without items keep the zeros written by the key pass.
__global__ __launch_bounds__(kBlock) void k_bucket_bounds(const uint64_t* __restrict__ items,
                                                          const Counters* __restrict__ counters,
                                                          uint32_t low,
                                                          uint32_t* __restrict__ bounds)
{
    const uint32_t entries = counters->grid_entries;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < entries;
         p += gridDim.x * blockDim.x) {
        const uint32_t b = uint32_t(items[p]) >> low;
        if (p == 0 || (uint32_t(items[p - 1]) >> low) != b)
            bounds[2 * b] = p;
        if (p + 1 == entries || (uint32_t(items[p + 1]) >> low) != b)
            bounds[2 * b + 1] = p + 1;
    }
}

// Records of the in-grid bodies in cell-key order, and every entry of the
// cell table; also the overflow list's length.  Workgroup `bucket` owns the keys whose bits [low, t) equal it,
// i.e. the 2^low cells bucket * 2^low + c, and the contiguous range [s, e) of
// the items sorted by those bits (single pass: s = the items of all smaller
// buckets, from the pass's digit totals; two passes: the range of the high
// digit from the totals, then the low digit's sub-range by search).  Counting
// sort in LDS: counts per cell, exclusive scan, the cells' table entries
// (start, end; start = end for an empty cell), then each record is copied to
// start + an LDS cursor (records of one cell in no particular order: the
// query's results do not depend on it).
template <uint32_t kLowMax>
__global__ __launch_bounds__(kBucketThreads) void k_grid_buckets(
    const uint64_t* __restrict__ items, uint32_t low, const uint32_t* __restrict__ bounds,
    const uint32_t* __restrict__ digit_total, const float4* __restrict__ aabb,
    float4* __restrict__ sorted, uint2* __restrict__ table, Counters* counters,
    uint32_t* __restrict__ over_ctr)
{
    __shared__ __align__(16) uint32_t cnt[1u << kLowMax];
    __shared__ uint32_t lds_wave[kBucketThreads / 64];
    const uint32_t bucket = blockIdx.x;
    if (bucket == 0 && threadIdx.x == 0) {
        // The key kernel's overflow list is complete: publish its length for
        // the query and the overflow pass, and zero the slot counter for the
        // next step.
        counters->overflow_count = *over_ctr;
        *over_ctr = 0;
    }
    const uint32_t cells = 1u << low;
    uint32_t s, e;
    if (bounds) {
        // Two passes: the range found by k_bucket_bounds (empty: 0, 0).
        s = bounds[2 * bucket];
        e = bounds[2 * bucket + 1];
    } else {
        // One pass: all items of smaller digits come before the bucket.
        const uint32_t d = threadIdx.x < 256 ? digit_total[threadIdx.x] : 0u;
        s = bucket_block_sum(threadIdx.x < bucket ? d : 0u, lds_wave);
        e = s + digit_total[bucket];
    }
    for (uint32_t c = threadIdx.x; c < cells; c += kBucketThreads)
        cnt[c] = 0;
    __syncthreads();
    const uint32_t cmask = cells - 1u;
    // kBucketItems items per thread and chunk, loads issued before their uses.
    for (uint32_t p0 = s; p0 < e; p0 += kBucketItems * kBucketThreads) {
        uint32_t c[kBucketItems];
#pragma unroll
        for (uint32_t r = 0; r < kBucketItems; ++r) {
            const uint32_t p = p0 + r * kBucketThreads + threadIdx.x;
            c[r] = uint32_t(items[p < e ? p : s]) & cmask;
        }
#pragma unroll
        for (uint32_t r = 0; r < kBucketItems; ++r)
            if (p0 + r * kBucketThreads + threadIdx.x < e)
                atomicAdd(&cnt[c[r]], 1u);
    }
    __syncthreads();
    // Exclusive scan of the counts: thread t takes cells [t * per, (t + 1) * per),
    // read and written four at a time.
    if (cells >= 4 * kBucketThreads) {
        const uint32_t per4 = cells / (4 * kBucketThreads);
        uint4* cnt4 = reinterpret_cast<uint4*>(cnt) + threadIdx.x * per4;
        uint32_t sum = 0;
        for (uint32_t k = 0; k < per4; ++k) {
            const uint4 v = cnt4[k];
            sum += v.x + v.y + v.z + v.w;
        }
        uint32_t run = bucket_block_scan(sum, lds_wave);
        __syncthreads();
        for (uint32_t k = 0; k < per4; ++k) {
            const uint4 v = cnt4[k];
            uint4 o;
            o.x = run;
            o.y = o.x + v.x;
            o.z = o.y + v.y;
            o.w = o.z + v.z;
            run = o.w + v.w;
            cnt4[k] = o;
        }
    } else {
        const uint32_t per = cells > kBucketThreads ? cells / kBucketThreads : 1u;
        const uint32_t c0 = threadIdx.x * per;
        uint32_t sum = 0;
        if (c0 < cells)
            for (uint32_t k = 0; k < per; ++k)
                sum += cnt[c0 + k];
        uint32_t run = bucket_block_scan(sum, lds_wave);
        __syncthreads();
        if (c0 < cells)
            for (uint32_t k = 0; k < per; ++k) {
                const uint32_t v = cnt[c0 + k];
                cnt[c0 + k] = run;
                run += v;
            }
    }
    __syncthreads();
    const uint32_t key0 = bucket << low;
    for (uint32_t c = threadIdx.x; c < cells; c += kBucketThreads) {
        const uint32_t end = c + 1 < cells ? cnt[c + 1] : e - s;
        table[key0 + c] = make_uint2(s + cnt[c], s + end);
    }
    __syncthreads();
    // Records are moved by lane pairs: both lanes of a pair take the same
    // item, the even one claims its position, and each moves one 16-byte half
    // of the record (one 32-byte access per record and instruction pair).
    const uint32_t q = threadIdx.x >> 1, half = threadIdx.x & 1u;
    constexpr uint32_t kPairs = kBucketThreads / 2;
    for (uint32_t p0 = s; p0 < e; p0 += 2 * kBucketItems * kPairs) {
        // Unconditional loads (an empty slot reads item s and record 0, and
        // stores nothing), so the arrays stay in registers.
        uint64_t it[2 * kBucketItems];
#pragma unroll
        for (uint32_t r = 0; r < 2 * kBucketItems; ++r) {
            const uint32_t p = p0 + r * kPairs + q;
            it[r] = items[p < e ? p : s];
        }
        uint32_t pos[2 * kBucketItems];
        float4 rec[2 * kBucketItems];
#pragma unroll
        for (uint32_t r = 0; r < 2 * kBucketItems; ++r) {
            const uint32_t p = p0 + r * kPairs + q;
            const size_t idx = p < e ? size_t(it[r] >> 32) : 0;
            rec[r] = aabb[2 * idx + half];
            uint32_t mine = 0;
            if (p < e && half == 0)
                mine = s + atomicAdd(&cnt[uint32_t(it[r]) & cmask], 1u);
            pos[r] = mine + (uint32_t)__shfl_xor((int)mine, 1);
        }
#pragma unroll
        for (uint32_t r = 0; r < 2 * kBucketItems; ++r) {
            const uint32_t p = p0 + r * kPairs + q;
            if (p < e)
                sorted[2 * size_t(pos[r]) + half] = rec[r];
        }
    }
}

// ---- Pair emission with the narrow phase's classification (collision step)
// The pairs a workgroup emits get their slots (emission indices) in
// pairs_raw as before, and, when `Classify::buckets` is set, also the work
// the narrow phase's classification does (physics_world.cpp:630-633, 670):
// hit[slot] = 0, and slot appended to the bucket of its shape-pair class
// unless both colliders belong to one body.  Counts are combined per
// wavefront flush (one atomic per class) and per workgroup at the end.
// Entries of a wavefront's pair queue in the grid query: 512 (4 KiB of LDS).
// Every mid-query flush takes a pair-counter and six class-counter atomics,
// all on one counter line; at 16M bodies (about 450 pairs per wavefront)
// 256 entries make those flushes cost the query a quarter of its time.
constexpr uint32_t kGridQueue = 512;

struct Classify {
    uint32_t* buckets;            // kClasses x cap candidate indices (null: plain emission)
    uint32_t* hit;
    uint32_t* class_count;        // Counters::class_count
    const uint32_t* body;         // collider bodies (null: one body per collider)
};

// Queue entries are pair records whose halves carry the colliders' shape
// types in their top two bits (the records' world-word bits); the emitted
// records have them cleared.
constexpr uint64_t kPairIndexBits = (uint64_t(kRecWorldMask) << 32) | kRecWorldMask;

__device__ __forceinline__ uint64_t typed_pair(uint32_t a, uint32_t world_a, uint32_t b,
                                               uint32_t world_b)
{
    const uint32_t ta = a | (world_a & ~kRecWorldMask), tb = b | (world_b & ~kRecWorldMask);
    return a < b ? (uint64_t(ta) << 32) | tb : (uint64_t(tb) << 32) | ta;
}

__device__ __forceinline__ uint32_t slot_class(const Classify& k, uint64_t v)
{
    const uint32_t hi32 = uint32_t(v >> 32), lo32 = uint32_t(v);
    if (k.body && k.body[hi32 & kRecWorldMask] == k.body[lo32 & kRecWorldMask])
        return kClasses;  // same body: erased before the narrow phase
    // shapePair (physics_world.cpp:670-671): SS, SB, BB, then the hull classes.
    const uint32_t ta = hi32 >> kRecTypeShift, tb = lo32 >> kRecTypeShift;
    const uint32_t lo = min(ta, tb), hi = max(ta, tb);
    return hi < 2 ? lo + hi : 3 + lo;
}

// Writes entries [0, count) of a wavefront's queue at slots base + i (those
// below cap), classifying them when k.buckets is set.  `cbase` (per class,
// wave-uniform) is where this wavefront's class entries start in the
// buckets; the per-class counts are what the caller reserved.  All lanes call
// it.
__device__ __forceinline__ void emit_classified(const uint64_t* buf, uint32_t count,
                                                unsigned long long base, uint64_t* __restrict__ out,
                                                uint64_t cap, const Classify& k,
                                                const uint32_t cbase[kClasses])
{
    const uint32_t lane = lane_id();
    uint32_t run[kClasses] = {0, 0, 0, 0, 0, 0};
    for (uint32_t i0 = 0; i0 < count; i0 += 64) {
        const uint32_t i = i0 + lane;
        const unsigned long long slot = base + i;
        const bool valid = i < count && slot < cap;
        uint32_t cls = kClasses;
        if (valid) {
            const uint64_t v = buf[i];
            out[slot] = v & kPairIndexBits;
            if (k.buckets) {
                k.hit[slot] = 0;
                cls = slot_class(k, v);
            }
        }
        if (!k.buckets)
            continue;
#pragma unroll
        for (uint32_t q = 0; q < kClasses; ++q) {
            const uint64_t m = __ballot(cls == q);
            if (cls == q)
                k.buckets[q * cap + cbase[q] + run[q] + (uint32_t)__popcll(m & lanes_below(lane))] =
                    uint32_t(slot);
            run[q] += (uint32_t)__popcll(m);
        }
    }
}

// Per-class counts of entries [0, count) of a wavefront's queue whose slots
// (base + i) are below cap.  All lanes call it; the counts are wave-uniform.
__device__ __forceinline__ void count_classes(const uint64_t* buf, uint32_t count,
                                              unsigned long long base, uint64_t cap,
                                              const Classify& k, uint32_t n[kClasses])
{
    const uint32_t lane = lane_id();
    for (uint32_t q = 0; q < kClasses; ++q)
        n[q] = 0;
    for (uint32_t i0 = 0; i0 < count; i0 += 64) {
        const uint32_t i = i0 + lane;
        const bool valid = i < count && base + i < cap;
        const uint32_t cls = valid ? slot_class(k, buf[i]) : kClasses;
#pragma unroll
        for (uint32_t q = 0; q < kClasses; ++q)
            n[q] += (uint32_t)__popcll(__ballot(cls == q));
    }
}

__device__ __forceinline__ void cq_flush(WaveQueue& q, uint64_t* __restrict__ out,
                                         unsigned long long* counter, uint64_t cap,
                                         const Classify& k)
{
    if (q.count == 0)
        return;
    wave_sync();
    const uint32_t lane = lane_id();
    unsigned long long base = 0;
    if (lane == 0)
        base = atomicAdd(counter, (unsigned long long)q.count);
    base = __shfl(base, 0);
    uint32_t cbase[kClasses] = {0, 0, 0, 0, 0, 0};
    if (k.buckets) {
        uint32_t n[kClasses];
        count_classes(q.buf, q.count, base, cap, k, n);
        uint32_t mine = 0;
#pragma unroll
        for (uint32_t c = 0; c < kClasses; ++c)
            if (lane == c && n[c])
                mine = atomicAdd(&k.class_count[c], n[c]);
#pragma unroll
        for (uint32_t c = 0; c < kClasses; ++c)
            cbase[c] = (uint32_t)__shfl((int)mine, c);
    }
    emit_classified(q.buf, q.count, base, out, cap, k, cbase);
    wave_sync();
    q.count = 0;
}

__device__ __forceinline__ void cq_push(WaveQueue& q, bool pred, uint64_t value,
                                        uint64_t* __restrict__ out, unsigned long long* counter,
                                        uint64_t cap, const Classify& k)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0)
        return;
    const uint32_t add = (uint32_t)__popcll(mask);
    if (q.count + add > kGridQueue)
        cq_flush(q, out, counter, cap, k);
    if (pred)
        q.buf[q.count + (uint32_t)__popcll(mask & lanes_below(lane_id()))] = value;
    q.count += add;
}

// Final flush of the workgroup's queues: one atomic for the pairs and one per
// class for the whole workgroup.  Every thread of the workgroup calls it.
__device__ __forceinline__ void cq_flush_block(WaveQueue& q, uint64_t* __restrict__ out,
                                               unsigned long long* counter, uint64_t cap,
                                               const Classify& k)
{
    __shared__ uint32_t offset[kMaxBlockWaves];
    __shared__ uint32_t class_off[kMaxBlockWaves][kClasses];
    __shared__ unsigned long long base_all;
    __shared__ uint32_t class_base[kClasses];
    const uint32_t wave = threadIdx.x >> 6, waves = (blockDim.x + 63) >> 6, lane = lane_id();
    __syncthreads();
    if (lane == 0)
        offset[wave] = q.count;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < waves; ++w) {
            const uint32_t c = offset[w];
            offset[w] = run;
            run += c;
        }
        base_all = run ? atomicAdd(counter, (unsigned long long)run) : 0ull;
    }
    __syncthreads();
    const unsigned long long base = base_all + offset[wave];
    uint32_t cbase[kClasses] = {0, 0, 0, 0, 0, 0};
    if (k.buckets) {
        uint32_t n[kClasses];
        count_classes(q.buf, q.count, base, cap, k, n);
        if (lane == 0)
            for (uint32_t c = 0; c < kClasses; ++c)
                class_off[wave][c] = n[c];
        __syncthreads();
        if (threadIdx.x < kClasses) {
            const uint32_t c = threadIdx.x;
            uint32_t run = 0;
            for (uint32_t w = 0; w < waves; ++w) {
                const uint32_t v = class_off[w][c];
                class_off[w][c] = run;
                run += v;
            }
            class_base[c] = run ? atomicAdd(&k.class_count[c], run) : 0u;
        }
        __syncthreads();
        for (uint32_t c = 0; c < kClasses; ++c)
            cbase[c] = class_base[c] + class_off[wave][c];
    }
    emit_classified(q.buf, q.count, base, out, cap, k, cbase);
    q.count = 0;
}

// Candidates of a body: bodies whose home cell lies, on every axis, in
// [home - 1, floor(max / cell)] (a partner's min is below this body's max and
// less than one cell below its min), i.e. 2 or 3 cells per axis.
//
// One wavefront takes 64 bodies (consecutive sorted positions) at a time and
// walks their candidates together: per z-slab each lane loads its up to nine
// cell ranges, a wave scan numbers all (body, candidate) tasks of the slab,
// and the tasks are tested 64 per round, one per lane: each round the owners
// write their tasks of the round (candidate position, owner lane, home-cell
// flag) into LDS and every lane reads one, loads that candidate's record and
// tests it against its owner's record (LDS).  The wave waits for
// ceil(tasks / 64) dependent record loads per slab instead of the largest
// candidate count of its lanes.  The neighbour keys are assembled from the
// spread coordinates of the three x, three y and two z cells.
//
// Bodies on the overflow path (rare) are tested here too, as the reference's
// overflow pass does (broadphase.cpp:162-166): every in-grid body against
// each of them, and each pair of them once (lower index first).
// Overflow bodies, run by the workgroups past the query's own (the query
// launch carries kOverflowGroups more): (o, j) for every in-grid body j, and
// (o, o') for two overflow bodies with o < o'.  Workgroup b of them takes
// overflow bodies b, b + kOverflowGroups, ...
constexpr uint32_t kOverflowGroups = 64;

__device__ __forceinline__ void grid_overflow_role(uint32_t entries, uint32_t over, uint32_t b,
                                                const float4* __restrict__ s,
                                                const float4* __restrict__ aabb,
                                                const uint32_t* __restrict__ overflow,
                                                Counters* counters, WaveQueue& queue,
                                                uint64_t* __restrict__ out, uint64_t cap,
                                                const Classify& cls)
{
    for (uint32_t k = b; k < over; k += kOverflowGroups) {
        const uint32_t o = overflow[k];
        const Rec ro = load_rec(aabb, o);
        for (uint32_t p0 = 0; p0 < entries; p0 += blockDim.x) {
            const uint32_t p = p0 + threadIdx.x;
            bool hit = false;
            uint32_t j = 0, wj = 0;
            if (p < entries) {
                const Rec rj = load_rec(s, p);
                j = rj.index;
                wj = rj.world;
                hit = rec_world(rj.world) == rec_world(ro.world) && rec_overlap(ro.b, rj.b);
            }
            cq_push(queue, hit, typed_pair(o, ro.world, j, wj), out, &counters->pair_count, cap,
                    cls);
        }
        for (uint32_t k0 = 0; k0 < over; k0 += blockDim.x) {
            const uint32_t k2 = k0 + threadIdx.x;
            bool hit = false;
            uint32_t j = 0, wj = 0;
            if (k2 < over) {
                j = overflow[k2];
                if (o < j) {
                    const Rec rj = load_rec(aabb, j);
                    wj = rj.world;
                    hit = rec_world(rj.world) == rec_world(ro.world) && rec_overlap(ro.b, rj.b);
                }
            }
            cq_push(queue, hit, typed_pair(o, ro.world, j, wj), out, &counters->pair_count, cap,
                    cls);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_grid_query(uint32_t n, const float4* __restrict__ s,
                                                       const uint32_t* __restrict__ table,
                                                       uint32_t bits_per_axis,
                                                       uint32_t table_bits, Counters* counters,
                                                       uint64_t* __restrict__ out, uint64_t cap,
                                                       const float4* __restrict__ aabb,
                                                       const uint32_t* __restrict__ overflow,
                                                       Classify cls)
{
    __shared__ uint64_t queue_lds[kBlock / 64][kGridQueue];
    __shared__ float4 own_lds[kBlock / 64][2][64];
    __shared__ uint32_t task_lds[kBlock / 64][64];
    __shared__ uint8_t owner_lds[kBlock / 64][64];
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    WaveQueue queue{queue_lds[wave], 0};
    const uint32_t entries = n - counters->overflow_count;
    const uint32_t groups = gridDim.x - kOverflowGroups;  // the query's own
    if (blockIdx.x >= groups) {
        grid_overflow_role(entries, counters->overflow_count, blockIdx.x - groups, s, aabb,
                           overflow, counters, queue, out, cap, cls);
        cq_flush_block(queue, out, &counters->pair_count, cap, cls);
        return;
    }
    float4* own = &own_lds[wave][0][0];
    uint32_t* task = task_lds[wave];
    uint8_t* task_owner = owner_lds[wave];
    const float cell = __uint_as_float(counters->cell_bits);
    const double inv = 1.0 / double(cell);
    const uint32_t mask = (1u << bits_per_axis) - 1u;
    const uint32_t tmask = (1u << table_bits) - 1u;
    const uint2* range = reinterpret_cast<const uint2*>(table);
    // Grid-stride over sorted positions; all lanes stay in the loops together
    // so that the wave-wide operations are executed convergently.
    const uint32_t stride = groups * blockDim.x;
    const uint32_t rounds = (entries + stride - 1) / stride;
    const uint32_t block = xcd_block(blockIdx.x, groups);
    for (uint32_t round = 0; round < rounds; ++round) {
        const uint32_t p = round * stride + block * blockDim.x + threadIdx.x;
        const bool active = p < entries;
        int32_t c[3] = {0, 0, 0}, top[3] = {0, 0, 0};  // both in range for in-grid bodies
        uint32_t wk = 0;
        if (active) {
            const float4 r0 = s[2 * size_t(p)], r1 = s[2 * size_t(p) + 1];
            own[lane] = r0;
            own[64 + lane] = r1;
            home_cell(r0.x, r0.y, r0.z, inv, c);
            home_cell(r0.w, r1.x, r1.y, inv, top);
            wk = world_key(rec_world(__float_as_uint(r1.w)), table_bits);
        }
        // Spread coordinates of the neighbour cells: x and y at -1, 0, +1,
        // z at 0, +1 (a cell's key is (sx | sy | sz) & tmask ^ wk).
        const uint32_t sx[3] = {spread3(uint32_t(c[0] - 1) & mask), spread3(uint32_t(c[0]) & mask),
                                spread3(uint32_t(c[0] + 1) & mask)};
        const uint32_t sy[3] = {spread3(uint32_t(c[1] - 1) & mask) << 1,
                                spread3(uint32_t(c[1]) & mask) << 1,
                                spread3(uint32_t(c[1] + 1) & mask) << 1};
        const uint32_t sz[2] = {spread3(uint32_t(c[2]) & mask) << 2,
                                spread3(uint32_t(c[2] + 1) & mask) << 2};
        // Half neighbourhood: of two overlapping bodies in different cells,
        // the one whose home cell comes first in (z, y, x) order finds the
        // pair (the other's cell is at offset (dx, dy, dz) > (0, 0, 0) in
        // that order, and within its [home - 1, top] range); in the home
        // cell itself the lower index does.  So the slab dz = -1 and the
        // first four cells of the slab dz = 0 are never read.
        for (int dz = 0; dz <= 1; ++dz) {
            const bool zok = active && c[2] + dz <= top[2];
            uint2 r[9];
            uint32_t cnt = 0;
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int dx = j % 3 - 1, dy = j / 3 - 1;
                r[j] = make_uint2(0u, 0u);
                if (zok && (dz > 0 || j >= 4) && c[0] + dx <= top[0] && c[1] + dy <= top[1])
                    r[j] = range[((sx[dx + 1] | sy[dy + 1] | sz[dz]) & tmask) ^ wk];
            }
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                r[j].y = r[j].y > r[j].x ? r[j].y - r[j].x : 0u;  // now (start, length)
                cnt += r[j].y;
            }
            // Exclusive prefix of the lanes' task counts, and their total.
            uint32_t incl = cnt;
#pragma unroll
            for (uint32_t off = 1; off < 64; off <<= 1) {
                const uint32_t t = (uint32_t)__shfl_up((int)incl, off);
                if (lane >= off)
                    incl += t;
            }
            const uint32_t first = incl - cnt;
            const uint32_t total = (uint32_t)__shfl((int)incl, 63);
#pragma unroll 1
            for (uint32_t base = 0; base < total; base += 64) {
                // Owners write their tasks in [base, base + 64).
                uint32_t so = first;
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    const uint32_t lo = max(so, base), hi = min(so + r[j].y, base + 64u);
                    const uint8_t tag = uint8_t(lane | ((dz == 0 && j == 4) ? 64u : 0u));
#pragma unroll 1
                    for (uint32_t t = lo; t < hi; ++t) {
                        task[t - base] = r[j].x + (t - so);
                        task_owner[t - base] = tag;
                    }
                    so += r[j].y;
                }
                wave_sync();
                bool hit = false;
                uint32_t a = 0, b = 0, wa = 0, wb = 0;
                if (base + lane < total) {
                    const uint32_t q = task[lane], w = task_owner[lane];
                    const uint32_t owner = w & 63u;
                    const bool home = (w >> 6) & 1u;
                    const float4 m0 = own[owner], m1 = own[64 + owner];
                    const Rec o = load_rec(s, q);
                    a = __float_as_uint(m1.z);
                    b = o.index;
                    wa = __float_as_uint(m1.w);
                    wb = o.world;
                    if ((!home || a < b) && rec_world(wa) == rec_world(wb)) {
                        const float mb[6] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y};
                        hit = rec_overlap(mb, o.b);
                    }
                }
                wave_sync();  // the next round rewrites the tasks
                cq_push(queue, hit, typed_pair(a, wa, b, wb), out, &counters->pair_count, cap,
                        cls);
            }
        }
        wave_sync();  // the next bodies rewrite the owner records
    }
    cq_flush_block(queue, out, &counters->pair_count, cap, cls);
}

}  // namespace

GridBits grid_bits(uint32_t n, uint32_t table_bits)
{
    GridBits g;
    g.table = std::min(grid_table_bits(n), table_bits);
    g.axis = g.table / 3;
    // A few thousand cells per bucket: up to 2^21 entries 8 bucket bits (one
    // pass), beyond one more cell bit per axis step (2^24: 14 + 10, 2^27:
    // 15 + 12, two passes).
    g.low = g.table <= 21 ? (g.table > 8 ? g.table - 8 : 0u) : std::min(13 + (g.table - 21) / 3, 15u);
    g.bucket = g.table - g.low;
    return g;
}
static_assert(grid_table_bits(uint64_t(1) << 30) - 15 <= 12, "bucket bits within kMaxGridBuckets");

void launch_grid(physgpu_context* ctx, double cell_size, bool fused)
{
    const uint32_t n = ctx->n;
    const GridBits g = grid_bits(n, ctx->table_bits);
    Timer t;
    if (fused) {
        launch_aabbs_grid_keys(ctx, float(cell_size), g);
    } else {
        time_begin(ctx, "grid_keys", t);
        const uint32_t tile_items = kSortThreads * ctx->tuning.sort_items;
        hipLaunchKernelGGL(k_grid_keys, dim3(div_up(n, tile_items)), dim3(kBlock), 0,
                           ctx->stream, n, ctx->aabb, ctx->counters, float(cell_size), g.axis,
                           g.table, g.low, ctx->items_a, ctx->overflow_list, ctx->grid_over_ctr,
                           tile_items, ctx->histogram, ctx->bucket_bounds,
                           g.bucket > 8 ? 2u << g.bucket : 0u);
        time_end(ctx, t);
    }
    // Items grouped by bucket (key bits [low, t)), overflow bodies dropped;
    // the kept count goes to grid_entries.
    uint64_t* grouped = nullptr;
    radix_sort(ctx, ctx->items_a, ctx->items_b, n, nullptr, g.low, g.bucket, 0, 0, &grouped,
               nullptr, nullptr, true, kOverflowKey, &ctx->counters->grid_entries);
    time_begin(ctx, "grid_buckets", t);
    const uint32_t* bounds = nullptr;
    if (g.bucket > 8) {
        hipLaunchKernelGGL(k_bucket_bounds, dim3(stride_grid(n)), dim3(kBlock), 0, ctx->stream,
                           grouped, ctx->counters, g.low, ctx->bucket_bounds);
        bounds = ctx->bucket_bounds;
    }
    const dim3 bg(1u << g.bucket), bb(kBucketThreads);
    uint2* table = reinterpret_cast<uint2*>(ctx->cell_table);
#define PG_BUCKETS(L)                                                                         \
    hipLaunchKernelGGL(k_grid_buckets<L>, bg, bb, 0, ctx->stream, grouped, g.low, bounds,      \
                       ctx->digit_total, ctx->aabb, ctx->sorted_aabb, table, ctx->counters,   \
                       ctx->grid_over_ctr)
    if (g.low <= 10)
        PG_BUCKETS(10);
    else if (g.low <= 13)
        PG_BUCKETS(13);
    else if (g.low <= 14)
        PG_BUCKETS(14);
    else
        PG_BUCKETS(15);
#undef PG_BUCKETS
    time_end(ctx, t);
    // The collision step's query also classifies the pairs for the narrow
    // phase (launch_narrowphase then skips its classification).
    Classify cls{nullptr, nullptr, nullptr, nullptr};
    if (fused) {
        cls = Classify{ctx->buckets, ctx->hit, ctx->counters->class_count,
                       ctx->has_world && ctx->world_has_body ? ctx->body : nullptr};
        ctx->np_classified = true;
    }
    time_begin(ctx, "grid_query", t);
    hipLaunchKernelGGL(k_grid_query, dim3(ctx->tuning.query_groups + kOverflowGroups),
                       dim3(ctx->tuning.query_block), 0, ctx->stream, n,
                       ctx->sorted_aabb, ctx->cell_table, g.axis, g.table, ctx->counters,
                       ctx->pairs_raw, ctx->cap_p, ctx->aabb, ctx->overflow_list, cls);
    time_end(ctx, t);
}

}  // namespace pg
