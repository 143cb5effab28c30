// Folding of per-workgroup partials into the step counters (see PartialSlot):
// one workgroup per entry reduces the slot's partials and writes the
// destination counter; no device atomics.
#include <stddef.h>

#include "internal.hpp"

namespace pg {
namespace {

constexpr uint32_t kFoldThreads = 1024;  // up to 8 partials per thread

__global__ __launch_bounds__(kFoldThreads) void k_fold(const uint64_t* __restrict__ partials,
                                                       FoldList list, Counters* counters)
{
    __shared__ uint64_t part[kFoldThreads / 64];
    if (list.finish_pairs && blockIdx.x == 0 && threadIdx.x == 0) {
        // launch_finish_pairs' work, after the broad phase's last pair.
        const unsigned long long p = counters->pair_count;
        counters->sort_count = uint32_t(p < list.finish_cap ? p : list.finish_cap);
        counters->mid_segments = 0;
        counters->long_segments = 0;
    }
    if (blockIdx.x >= list.count)
        return;
    const FoldEntry e = list.e[blockIdx.x];
    const uint64_t* src = partials + size_t(e.slot) * kMaxPartialBlocks;
    const bool is_min = e.op == kFoldMin32, is_max = e.op == kFoldMax32;
    uint64_t v = is_min ? 0xFFFFFFFFull : 0ull;
    for (uint32_t b = threadIdx.x; b < e.blocks; b += blockDim.x) {
        const uint64_t x = src[b];
        v = is_min ? (x < v ? x : v) : (is_max ? (x > v ? x : v) : v + x);
    }
    for (uint32_t off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off);
        v = is_min ? (o < v ? o : v) : (is_max ? (o > v ? o : v) : v + o);
    }
    if ((threadIdx.x & 63u) == 0)
        part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
        const uint64_t o = part[w];
        v = is_min ? (o < v ? o : v) : (is_max ? (o > v ? o : v) : v + o);
    }
    char* dst = reinterpret_cast<char*>(counters) + e.offset;
    if (e.op == kFoldSum64) {
        *reinterpret_cast<unsigned long long*>(dst) += v;
    } else {
        uint32_t* d = reinterpret_cast<uint32_t*>(dst);
        const uint32_t x = uint32_t(v);
        *d = e.op == kFoldSum32 ? *d + x : (is_max ? (x > *d ? x : *d) : (x < *d ? x : *d));
    }
}

}  // namespace

void launch_fold(physgpu_context* ctx, const FoldList& in)
{
    FoldList list = in;
    if (ctx->finish_pending) {
        list.finish_pairs = 1;
        list.finish_cap = ctx->cap_p;
        ctx->finish_pending = false;
    }
    if (list.count || list.finish_pairs)
        hipLaunchKernelGGL(k_fold, dim3(list.count ? list.count : 1), dim3(kFoldThreads), 0, ctx->stream,
                           ctx->partials, list, ctx->counters);
}

}  // namespace pg
