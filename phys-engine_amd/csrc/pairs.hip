// Canonical pair order (src/collision/broadphase.cpp:18-60).
//
// Every broad phase appends unordered (first, second) records, first < second,
// as first << 32 | second.  The reference orders them lexicographically with
// two stable counting passes, by `second` then by `first` (sortPairsLinear).
// Here the same two stable passes are the LSD radix sort over the bits of
// `second` and then of `first` (ceil(log2 N) bits each), reading the pair
// count from the device.
#include "internal.hpp"

namespace pg {
namespace {

__global__ void k_finish_pairs(Counters* c, uint64_t cap)
{
    const unsigned long long p = c->pair_count;
    c->sort_count = uint32_t(p < cap ? p : cap);
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

uint32_t index_bits(uint32_t n)
{
    uint32_t b = 1;
    while (b < 32 && (1ull << b) < n)
        ++b;
    return b;
}

}  // namespace

void launch_canonical_pairs(physgpu_context* ctx)
{
    hipLaunchKernelGGL(k_finish_pairs, dim3(1), dim3(1), 0, ctx->stream, ctx->counters,
                       ctx->cap_p);
    const uint32_t bits = index_bits(ctx->n);
    uint64_t* sorted = nullptr;
    radix_sort(ctx, ctx->pairs_raw, ctx->pairs_tmp, uint32_t(ctx->cap_p),
               &ctx->counters->sort_count, 0, bits, 32, bits, &sorted);
    ctx->pairs = sorted;
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
