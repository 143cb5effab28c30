// Folding of one slot of per-workgroup partials (see PartialSlot) into its
// step counter, by one workgroup of kFoldThreads threads: shared by k_fold and
// the kernels that fold in their spare workgroups.
#pragma once

#include <stddef.h>

#include "internal.hpp"

namespace pg {

constexpr uint32_t kFoldThreads = 1024;  // up to 8 partials per thread

// Every thread of the (kFoldThreads-thread) workgroup calls it.
__device__ __forceinline__ void fold_entry(const uint64_t* __restrict__ partials,
                                           const FoldEntry& e, Counters* counters)
{
    __shared__ uint64_t part[kFoldThreads / 64];
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

}  // namespace pg
