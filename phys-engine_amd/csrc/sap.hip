// Sweep and prune on the x axis (src/collision/broadphase.cpp:205-291).
//
// Records are sorted by (min.x, index) with the radix sort, then gathered into
// a sorted copy of the 32-byte AABB records.  The reference scans, for each row i, the records j > i
// until min_j.x > max_i.x and emits the pair when Y and Z overlap inclusively.
// On gfx950 the sweep is turned around: a workgroup owns 4096 consecutive
// sorted records as *columns* (4 per lane, held in registers) and streams the
// rows that can reach them through LDS.  A row i can reach column j only when
// min_i.x >= min_j.x - max side, so the candidate rows of a workgroup form a
// contiguous range found by one binary search.  The workgroup's columns are
// ordered by min.y in LDS and dealt to its wavefronts in that order, so each
// wavefront holds a band of y.  The thread that stages a row in LDS lists it
// for every wavefront whose columns its window and y interval can reach; each
// wavefront reads only its listed rows (broadcast LDS reads), and all 64
// lanes test their 4 columns.  Hits are compacted with a wavefront ballot and
// one atomic per wave.
#include <stddef.h>

#include "devmath.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace pg {
namespace {

constexpr uint32_t kColsPerLane = 4;
constexpr uint32_t kColsPerWave = 64 * kColsPerLane;
// 1024 threads: 4096 columns and sixteen y bands per workgroup (at 1M bodies
// 256 threads sweep in 0.63 ms, 512 in 0.43, 1024 in 0.33; about 120 KB of
// LDS, one workgroup per CU).
constexpr uint32_t kSweepThreads = 1024;
constexpr uint32_t kColsPerBlock = kColsPerWave * (kSweepThreads / 64);
constexpr uint32_t kRowChunk = kSweepThreads;

__global__ __launch_bounds__(kBlock) void k_sap_keys(uint32_t n, const float4* __restrict__ aabb,
                                                     uint64_t* __restrict__ items)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    items[i] = (uint64_t)float_key(aabb[2 * size_t(i)].x) | ((uint64_t)i << 32);
}

// Sorted copy of the records (the index rides in the record).
__global__ __launch_bounds__(kBlock) void k_gather_sorted(uint32_t n,
                                                          const uint64_t* __restrict__ items,
                                                          const float4* __restrict__ aabb,
                                                          float4* __restrict__ sorted)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const size_t idx = size_t(items[i] >> 32);
    sorted[2 * size_t(i)] = aabb[2 * idx];
    sorted[2 * size_t(i) + 1] = aabb[2 * idx + 1];
}

__global__ __launch_bounds__(kSweepThreads) void k_sap_sweep(uint32_t n, const float4* __restrict__ s,
                                                      const Counters* counters_ro,
                                                      Counters* counters,
                                                      uint64_t* __restrict__ out, uint64_t cap)
{
    // Rows staged in LDS as {max.x, min.y, max.y, min.z} and {max.z, index}.
    __shared__ float4 row_a[kRowChunk];
    __shared__ float2 row_b[kRowChunk];
    __shared__ uint32_t s_lo;
    __shared__ uint64_t queue_lds[kSweepThreads / 64][kWaveQueue];
    WaveQueue queue{queue_lds[threadIdx.x >> 6], 0};

    __shared__ uint64_t order[kColsPerBlock];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t j0 = blockIdx.x * kColsPerBlock;

    // The workgroup's columns ordered by min.y (bitonic sort of (min.y key,
    // column) in LDS), so that each wavefront holds a band of y: a row whose
    // y interval misses a wavefront's band skips it with a uniform branch.
    for (uint32_t c = threadIdx.x; c < kColsPerBlock; c += kSweepThreads) {
        const uint32_t j = j0 + c;
        const uint32_t key = j < n ? float_key(s[2 * size_t(j)].y) : 0xFFFFFFFFu;
        order[c] = (uint64_t(key) << 32) | c;
    }
    __syncthreads();
    for (uint32_t size = 2; size <= kColsPerBlock; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < kColsPerBlock / 2; t += kSweepThreads) {
                const uint32_t lo = 2 * t - (t & (stride - 1));
                const uint32_t hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t a = order[lo], b = order[hi];
                if ((a > b) == up) {
                    order[lo] = b;
                    order[hi] = a;
                }
            }
            __syncthreads();
        }
    }

    // Columns in registers: wavefront w takes sorted positions
    // [256 w, 256 w + 256).
    float c_minx[kColsPerLane], c_miny[kColsPerLane], c_maxy[kColsPerLane];
    float c_minz[kColsPerLane], c_maxz[kColsPerLane];
    uint32_t c_idx[kColsPerLane], c_j[kColsPerLane];
    const uint32_t wave = threadIdx.x >> 6;
    float wave_min_x = __int_as_float(0x7f800000), ylo = __int_as_float(0x7f800000);
    float yhi = __int_as_float(0xff800000);
    uint32_t wave_last = 0;  // rows must be below this wavefront's last column
#pragma unroll
    for (uint32_t k = 0; k < kColsPerLane; ++k) {
        const uint32_t j = j0 + uint32_t(order[wave * kColsPerWave + k * 64u + lane] & 0xFFFFFFFFu);
        c_j[k] = j;
        if (j < n) {
            const Rec r = load_rec(s, j);
            c_minx[k] = r.b[0];
            c_miny[k] = r.b[1];
            c_minz[k] = r.b[2];
            c_maxy[k] = r.b[4];
            c_maxz[k] = r.b[5];
            c_idx[k] = r.index;
            wave_min_x = fminf(wave_min_x, r.b[0]);
            ylo = fminf(ylo, r.b[1]);
            yhi = fmaxf(yhi, r.b[4]);
            wave_last = max(wave_last, j + 1);
        } else {
            c_minx[k] = __int_as_float(0x7f800000);  // +inf: never <= max_i.x
            c_miny[k] = c_maxy[k] = c_minz[k] = c_maxz[k] = 0.0f;
            c_idx[k] = 0;
        }
    }
    for (uint32_t off = 32; off > 0; off >>= 1) {
        wave_min_x = fminf(wave_min_x, __shfl_xor(wave_min_x, off));
        ylo = fminf(ylo, __shfl_xor(ylo, off));
        yhi = fmaxf(yhi, __shfl_xor(yhi, off));
        wave_last = max(wave_last, (uint32_t)__shfl_xor((int)wave_last, off));
    }
    // (fminf / fmaxf pass over a NaN bound: such a column meets no row.)

    // Each wavefront's row filter, shared: a row of a chunk goes onto the
    // list of every wavefront whose window and y band it reaches.
    constexpr uint32_t kWaves = kSweepThreads / 64;
    __shared__ float band_lds[kWaves][3];
    __shared__ uint32_t last_lds[kWaves];
    __shared__ uint32_t list_count[2][kWaves];  // by chunk parity
    __shared__ uint16_t list_rows[kWaves][kRowChunk];
    if (threadIdx.x < 2 * kWaves)
        list_count[threadIdx.x / kWaves][threadIdx.x % kWaves] = 0;
    if (lane == 0) {
        band_lds[wave][0] = wave_min_x;
        band_lds[wave][1] = ylo;
        band_lds[wave][2] = yhi;
        last_lds[wave] = wave_last;
    }

    if (threadIdx.x == 0) {
        // First row that can reach column j0: min_i.x >= min_j0.x - max side.
        const float wmax = key_float(counters_ro->wmax_x_key);
        const float x0 = s[2 * size_t(j0)].x;
        const float bound = (x0 - wmax * 1.001f) - fabsf(x0) * 1e-6f;
        uint32_t lo = 0, hi = j0;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s[2 * size_t(mid)].x < bound)
                lo = mid + 1;
            else
                hi = mid;
        }
        s_lo = lo;
    }
    __syncthreads();
    const uint32_t row_begin = s_lo;
    const uint32_t row_end = min(j0 + kColsPerBlock, n) - 1;  // rows < last column

    uint32_t parity = 0;
    for (uint32_t chunk = row_begin; chunk < row_end; chunk += kRowChunk, parity ^= 1u) {
        // Every wavefront has read the previous chunk's lists and counts.
        __syncthreads();
        {
            const uint32_t i = chunk + threadIdx.x;
            if (i < row_end) {
                const Rec r = load_rec(s, i);
                row_a[threadIdx.x] = make_float4(r.b[3], r.b[1], r.b[4], r.b[2]);
                row_b[threadIdx.x] = make_float2(r.b[5], __uint_as_float(r.index));
                // The wavefronts this row can meet: its window reaches their
                // columns (max.x >= their least min.x, a column after it) and
                // its y interval meets their band.
#pragma unroll
                for (uint32_t w = 0; w < kWaves; ++w) {
                    if (r.b[3] >= band_lds[w][0] && r.b[4] >= band_lds[w][1] &&
                        r.b[1] <= band_lds[w][2] && i + 1 < last_lds[w])
                        list_rows[w][atomicAdd(&list_count[parity][w], 1u)] =
                            uint16_t(threadIdx.x);
                }
            }
        }
        __syncthreads();
        const uint32_t rows = list_count[parity][wave];
        // The other parity's counts were read before this chunk's first
        // barrier; they start the next chunk at zero.
        if (threadIdx.x < kWaves)
            list_count[parity ^ 1u][threadIdx.x] = 0;
        // Four rows per step: their LDS reads are issued together.
        for (uint32_t r0 = 0; r0 < rows; r0 += 4) {
            float4 ra[4];
            float2 rb[4];
            uint32_t ro[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                ro[u] = list_rows[wave][min(r0 + u, rows - 1)];
                ra[u] = row_a[ro[u]];
                rb[u] = row_b[ro[u]];
            }
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                if (r0 + u >= rows)
                    continue;
                const uint32_t i = chunk + ro[u];
                const float rx = ra[u].x, ry0 = ra[u].y, ry1 = ra[u].z, rz0 = ra[u].w;
                const float rz1 = rb[u].x;
                const uint32_t ri = __float_as_uint(rb[u].y);
                bool hit[kColsPerLane];
                bool any = false;
#pragma unroll
                for (uint32_t k = 0; k < kColsPerLane; ++k) {
                    hit[k] = c_j[k] > i && c_minx[k] <= rx && ry0 <= c_maxy[k] &&
                             ry1 >= c_miny[k] && rz0 <= c_maxz[k] && rz1 >= c_minz[k];
                    any = any || hit[k];
                }
                // One ballot per row; the queue is touched only on a hit.
                if (__ballot(any) == 0)
                    continue;
#pragma unroll
                for (uint32_t k = 0; k < kColsPerLane; ++k)
                    wq_push(queue, hit[k], pair_value(ri, c_idx[k]), out, &counters->pair_count,
                            cap);
            }
        }
    }
    wq_flush_block(queue, out, &counters->pair_count, cap);
}

// xWindowComparisons (broadphase.cpp:241): for row i, the reference tests
// j = i+1 .. first j with min_j.x > max_i.x, counting j - i - 1.
__global__ __launch_bounds__(kBlock) void k_window_tests(uint32_t n, const float4* __restrict__ s,
                                                         uint64_t* __restrict__ tests_part)
{
    unsigned long long total = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        const float x = s[2 * size_t(i)].w;
        uint32_t lo = i + 1, hi = n;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s[2 * size_t(mid)].x > x)
                hi = mid;
            else
                lo = mid + 1;
        }
        total += lo - i - 1;
    }
    block_partial(total, tests_part);
}

}  // namespace

void launch_sap_keys(physgpu_context* ctx)
{
    Timer t;
    time_begin(ctx, "sap_keys", t);
    hipLaunchKernelGGL(k_sap_keys, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->n, ctx->aabb, ctx->items_a);
    time_end(ctx, t);
}

void launch_gather_sorted(physgpu_context* ctx, const uint64_t* sorted_items)
{
    Timer t;
    time_begin(ctx, "gather_sorted", t);
    hipLaunchKernelGGL(k_gather_sorted, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0,
                       ctx->stream, ctx->n, sorted_items, ctx->aabb, ctx->sorted_aabb);
    time_end(ctx, t);
}

void launch_sap_sweep(physgpu_context* ctx)
{
    Timer t;
    time_begin(ctx, "sap_sweep", t);
    hipLaunchKernelGGL(k_sap_sweep, dim3(div_up(ctx->n, kColsPerBlock)), dim3(kSweepThreads), 0,
                       ctx->stream, ctx->n, ctx->sorted_aabb, ctx->counters, ctx->counters,
                       ctx->pairs_raw, ctx->cap_p);
    time_end(ctx, t);
}

void launch_sap_window_tests(physgpu_context* ctx)
{
    Timer t;
    time_begin(ctx, "sap_window_tests", t);
    const uint32_t grid = stride_grid(ctx->n);
    hipLaunchKernelGGL(k_window_tests, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->n,
                       ctx->sorted_aabb, ctx->partials + size_t(kSlotWindowTests) * kMaxPartialBlocks);
    time_end(ctx, t);
    FoldList fold{};
    fold.e[0] = {kSlotWindowTests, grid, kFoldSum64, uint32_t(offsetof(Counters, window_tests))};
    fold.count = 1;
    launch_fold(ctx, fold);
}

}  // namespace pg
