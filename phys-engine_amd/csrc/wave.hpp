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

namespace pg {

// AABB record in HBM: two float4, {min.x, min.y, min.z, max.x} and
// {max.y, max.z, index bits, 0}; 32 bytes, one sector, so a scattered
// candidate costs one 32-byte access instead of six 4-byte planes.
struct Rec {
    float b[6];  // min x y z, max x y z
    uint32_t index;
};

__device__ __forceinline__ Rec load_rec(const float4* __restrict__ r, size_t i)
{
    const float4 a = r[2 * i], c = r[2 * i + 1];
    return {{a.x, a.y, a.z, a.w, c.x, c.y}, __float_as_uint(c.z)};
}

__device__ __forceinline__ void store_rec(float4* __restrict__ r, size_t i, const float b[6],
                                          uint32_t index)
{
    r[2 * i] = make_float4(b[0], b[1], b[2], b[3]);
    r[2 * i + 1] = make_float4(b[4], b[5], __uint_as_float(index), 0.0f);
}

// Inclusive AABB overlap (Aabb::overlaps, include/phys/collision/aabb.hpp:15-19).
__device__ __forceinline__ bool rec_overlap(const float a[6], const float b[6])
{
    return a[0] <= b[3] && a[3] >= b[0] && a[1] <= b[4] && a[4] >= b[1] && a[2] <= b[5] &&
           a[5] >= b[2];
}

}  // namespace pg

namespace pg {

// Orders LDS traffic of one wavefront: every lane's earlier LDS accesses are
// complete and visible before any lane's later ones.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wavefront pair queue in LDS.  Hits are staged here and written out
// kWaveQueue at a time with one global atomic, instead of one atomic per
// ballot (a single contended counter serialises otherwise).  `count` is
// wave-uniform.  All lanes must call push/flush together.
constexpr uint32_t kWaveQueue = 256;

struct WaveQueue {
    uint64_t* buf;  // kWaveQueue entries of this wave's LDS
    uint32_t count;
};

__device__ __forceinline__ void wq_flush(WaveQueue& q, uint64_t* __restrict__ out,
                                         unsigned long long* counter, uint64_t cap)
{
    if (q.count == 0)
        return;
    wave_sync();
    const uint32_t lane = lane_id();
    unsigned long long base = 0;
    if (lane == 0)
        base = atomicAdd(counter, (unsigned long long)q.count);
    base = __shfl(base, 0);
    for (uint32_t i = lane; i < q.count; i += 64) {
        const unsigned long long slot = base + i;
        if (slot < cap)
            out[slot] = q.buf[i];
    }
    wave_sync();
    q.count = 0;
}

__device__ __forceinline__ void wq_push(WaveQueue& q, bool pred, uint64_t value,
                                        uint64_t* __restrict__ out, unsigned long long* counter,
                                        uint64_t cap)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0)
        return;
    const uint32_t add = (uint32_t)__popcll(mask);
    if (q.count + add > kWaveQueue)
        wq_flush(q, out, counter, cap);
    if (pred)
        q.buf[q.count + (uint32_t)__popcll(mask & lanes_below(lane_id()))] = value;
    q.count += add;
}

}  // namespace pg
