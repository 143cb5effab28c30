// Wavefront (64-lane) helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint64_t lanes_below(uint32_t lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

// Inverse of float_key (devmath.hpp).
__device__ __forceinline__ float key_float(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

// Wave-aggregated append: lanes with pred write one u64 each to out at
// consecutive slots of a global counter, one atomic per wave.  Returns
// nothing; slots at or past cap are dropped (the counter still counts them,
// which is how capacity overflow is reported).  All lanes must call it.
__device__ __forceinline__ void wave_append(bool pred, uint64_t value, uint64_t* __restrict__ out,
                                            unsigned long long* counter, uint64_t cap)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0)
        return;
    const uint32_t lane = lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    unsigned long long base = 0;
    if (lane == leader)
        base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, (int)leader);
    if (pred) {
        const unsigned long long slot = base + (unsigned long long)__popcll(mask & lanes_below(lane));
        if (slot < cap)
            out[slot] = value;
    }
}

// Pair record: first << 32 | second with first < second.
__device__ __forceinline__ uint64_t pair_value(uint32_t a, uint32_t b)
{
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return ((uint64_t)lo << 32) | hi;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
    for (uint32_t off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    return v;
}

}  // namespace pg
