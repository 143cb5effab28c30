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

// MI355X dispatches workgroups round-robin over its 8 XCDs, each with its own
// L2.  Renumbering workgroup b of `groups` so that each XCD runs one
// contiguous range of logical workgroups keeps neighbouring (spatially
// sorted) work in the same L2.  Identity when groups is not a multiple of 8.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t groups)
{
    constexpr uint32_t kXcds = 8;
    if (groups % kXcds)
        return b;
    return (b % kXcds) * (groups / kXcds) + b / kXcds;
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
// {max.y, max.z, index bits, world bits}; 32 bytes, one sector, so a
// scattered candidate costs one 32-byte access instead of six 4-byte planes.
// `world` is the collider's world in a batch of worlds (0 for one world):
// only records of the same world can pair.  Records written by the uniform
// grid's key pass also carry the collider's shape type in the world word's
// top two bits (colliders and worlds number at most 2^30), so that the grid
// query classifies the pairs it emits without reading the shape types.
constexpr uint32_t kRecTypeShift = 30;
constexpr uint32_t kRecWorldMask = (1u << kRecTypeShift) - 1u;

__device__ __forceinline__ uint32_t rec_world(uint32_t world_word)
{
    return world_word & kRecWorldMask;
}

struct Rec {
    float b[6];  // min x y z, max x y z
    uint32_t index;
    uint32_t world;
};

__device__ __forceinline__ Rec load_rec(const float4* __restrict__ r, size_t i)
{
    const float4 a = r[2 * i], c = r[2 * i + 1];
    return {{a.x, a.y, a.z, a.w, c.x, c.y}, __float_as_uint(c.z), __float_as_uint(c.w)};
}

__device__ __forceinline__ void store_rec(float4* __restrict__ r, size_t i, const float b[6],
                                          uint32_t index, uint32_t world = 0)
{
    r[2 * i] = make_float4(b[0], b[1], b[2], b[3]);
    r[2 * i + 1] = make_float4(b[4], b[5], __uint_as_float(index), __uint_as_float(world));
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

// Device-scope atomics on one address are serialised by the memory system
// (about 10 ns each on MI355X), so per-wavefront atomics at the end of a
// kernel with thousands of wavefronts form a tail of their own.  These
// helpers combine a workgroup first: one atomic per workgroup.  Every thread
// of the workgroup must call them (they synchronise the workgroup).
constexpr uint32_t kMaxBlockWaves = 16;

template <typename T>
__device__ __forceinline__ void block_add(T v, T* counter)
{
    __shared__ T part[kMaxBlockWaves];
    for (uint32_t off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    const uint32_t wave = threadIdx.x >> 6, waves = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane_id() == 0)
        part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T sum = 0;
        for (uint32_t w = 0; w < waves; ++w)
            sum += part[w];
        if (sum)
            atomicAdd(counter, sum);
    }
}

// Workgroup sum of v written (by thread 0, no atomic) to slot[blockIdx.x].
// Every thread of the workgroup must call it.
template <typename T>
__device__ __forceinline__ void block_partial(T v, uint64_t* slot)
{
    __shared__ uint64_t part[kMaxBlockWaves];
    uint64_t x = uint64_t(v);
    for (uint32_t off = 32; off > 0; off >>= 1)
        x += __shfl_xor(x, off);
    const uint32_t wave = threadIdx.x >> 6, waves = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane_id() == 0)
        part[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t sum = 0;
        for (uint32_t w = 0; w < waves; ++w)
            sum += part[w];
        slot[blockIdx.x] = sum;
    }
}

// Final flush of every wavefront's queue of the workgroup: one atomic for the
// workgroup's total, then each wavefront writes its entries at its offset.
__device__ __forceinline__ void wq_flush_block(WaveQueue& q, uint64_t* __restrict__ out,
                                               unsigned long long* counter, uint64_t cap)
{
    __shared__ uint32_t offset[kMaxBlockWaves];
    __shared__ unsigned long long base_all;
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
    for (uint32_t i = lane; i < q.count; i += 64) {
        const unsigned long long slot = base + i;
        if (slot < cap)
            out[slot] = q.buf[i];
    }
    q.count = 0;
}

}  // namespace pg

namespace pg {

// Cross-lane moves through DPP (data-parallel primitives: a VALU operand
// modifier, no LDS round trip).  Controls used: quad_perm [1,0,3,2] (0xB1),
// quad_perm [2,3,0,1] (0x4E), row_half_mirror (0x141), row_mirror (0x140);
// every lane of the row has a source, so every lane is written.
template <int Ctrl>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, Ctrl, 0xF, 0xF, false);
}

// Value of lane `src` (wave-uniform index) in every lane.
__device__ __forceinline__ uint32_t lane_value(uint32_t v, uint32_t src)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)src);
}

__device__ __forceinline__ float lane_value(float v, uint32_t src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)src));
}

__device__ __forceinline__ uint32_t uniform(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

constexpr uint32_t kNoIndex = 0xFFFFFFFFu;

// (value, index) candidates; index kNoIndex marks an empty lane.  Max=false
// keeps the smallest value, Max=true the largest; equal values keep the lower
// index.  This is a total order on the non-empty candidates (NaN aside), so
// the reduction tree does not change the result: it is the first extreme
// of a sequential scan in index order.
template <bool Max>
__device__ __forceinline__ void arg_take(float& v, uint32_t& i, float ov, uint32_t oi)
{
    const bool better = Max ? ov > v : ov < v;
    if (oi != kNoIndex && (i == kNoIndex || better || (ov == v && oi < i))) {
        v = ov;
        i = oi;
    }
}

template <bool Max, int Ctrl>
__device__ __forceinline__ void arg_step(float& v, uint32_t& i)
{
    const float ov = __uint_as_float(dpp_u32<Ctrl>(__float_as_uint(v)));
    const uint32_t oi = dpp_u32<Ctrl>(i);
    arg_take<Max>(v, i, ov, oi);
}

// Arg-min / arg-max within each 16-lane row (the DPP row): every lane of a
// row ends with its row's result.
template <bool Max>
__device__ __forceinline__ void row_arg_best(float& v, uint32_t& i)
{
    arg_step<Max, 0xB1>(v, i);
    arg_step<Max, 0x4E>(v, i);
    arg_step<Max, 0x141>(v, i);
    arg_step<Max, 0x140>(v, i);
}

// Wave-wide arg-min / arg-max; the result is returned uniform in every lane.
template <bool Max>
__device__ __forceinline__ void wave_arg_best(float& v, uint32_t& i)
{
    arg_step<Max, 0xB1>(v, i);
    arg_step<Max, 0x4E>(v, i);
    arg_step<Max, 0x141>(v, i);
    arg_step<Max, 0x140>(v, i);
    // Every lane of a 16-lane row now holds the row's result.
    float bv = lane_value(v, 0);
    uint32_t bi = lane_value(i, 0);
    arg_take<Max>(bv, bi, lane_value(v, 16), lane_value(i, 16));
    arg_take<Max>(bv, bi, lane_value(v, 32), lane_value(i, 32));
    arg_take<Max>(bv, bi, lane_value(v, 48), lane_value(i, 48));
    v = bv;
    i = bi;
}

template <int Ctrl>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, 0xF, 0xF, false));
}

// wave_arg_best for candidates whose index is chunk * 64 + lane (each lane
// holding the first extreme of its own chunks, scanned in order): the
// extreme value by DPP min / max, then the lowest chunk and lane holding it
// by ballots, over `chunks` chunks.  Returns false, leaving v and i as they
// were, when a candidate value is NaN (the caller then runs wave_arg_best,
// whose strict comparisons give NaN the sequential scan's treatment).
template <bool Max>
__device__ __forceinline__ bool wave_arg_best_lanes(float& v, uint32_t& i, uint32_t chunks)
{
    const bool has = i != kNoIndex;
    if (__ballot(has && v != v))
        return false;
    const float none = Max ? -__builtin_inff() : __builtin_inff();
    float x = has ? v : none;
    auto pick = [](float a, float b) { return Max ? fmaxf(a, b) : fminf(a, b); };
    x = pick(x, dpp_f32<0xB1>(x));
    x = pick(x, dpp_f32<0x4E>(x));
    x = pick(x, dpp_f32<0x141>(x));
    x = pick(x, dpp_f32<0x140>(x));
    // Rows to wave: v_permlane16_swap pairs rows 0-1 and 2-3, then
    // v_permlane32_swap the two halves; both results of each swap hold the
    // two partners' values, so every lane ends with the wave's extreme.
    const auto r16 = __builtin_amdgcn_permlane16_swap(__float_as_int(x), __float_as_int(x), false,
                                                      false);
    x = pick(__int_as_float(r16[0]), __int_as_float(r16[1]));
    const auto r32 = __builtin_amdgcn_permlane32_swap(__float_as_int(x), __float_as_int(x), false,
                                                      false);
    const float m = pick(__int_as_float(r32[0]), __int_as_float(r32[1]));
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint64_t at = __ballot(has && v == m && (i >> 6) == c);
        if (at) {
            i = c * 64 + (uint32_t)__ffsll((unsigned long long)at) - 1u;
            v = m;
            return true;
        }
    }
    return false;
}

// First minimum over the wavefront of (v, i) pairs (lanes with i == kNoIndex
// hold none): the lowest v by DPP min, then the lowest index holding it by
// DPP min, so ties go to the lower index as in a sequential strict-< scan.
// Every lane ends with the result.  A NaN v is passed over (callers that can
// meet one test for it themselves).
__device__ __forceinline__ void wave_min_first(float& v, uint32_t& i)
{
    float x = i != kNoIndex && v == v ? v : __builtin_inff();
    x = fminf(x, dpp_f32<0xB1>(x));
    x = fminf(x, dpp_f32<0x4E>(x));
    x = fminf(x, dpp_f32<0x141>(x));
    x = fminf(x, dpp_f32<0x140>(x));
    const auto x16 = __builtin_amdgcn_permlane16_swap(__float_as_int(x), __float_as_int(x), false,
                                                      false);
    x = fminf(__int_as_float(x16[0]), __int_as_float(x16[1]));
    const auto x32 = __builtin_amdgcn_permlane32_swap(__float_as_int(x), __float_as_int(x), false,
                                                      false);
    const float m = fminf(__int_as_float(x32[0]), __int_as_float(x32[1]));
    uint32_t k = i != kNoIndex && v == m ? i : kNoIndex;
    k = min(k, dpp_u32<0xB1>(k));
    k = min(k, dpp_u32<0x4E>(k));
    k = min(k, dpp_u32<0x141>(k));
    k = min(k, dpp_u32<0x140>(k));
    const auto k16 = __builtin_amdgcn_permlane16_swap((int)k, (int)k, false, false);
    k = min((uint32_t)k16[0], (uint32_t)k16[1]);
    const auto k32 = __builtin_amdgcn_permlane32_swap((int)k, (int)k, false, false);
    k = min((uint32_t)k32[0], (uint32_t)k32[1]);
    v = m;
    i = k;
}

// row_arg_best for candidates whose index is chunk * 16 + (lane within the
// 16-lane row), each lane holding the first extreme of its own chunks: the
// row's extreme by DPP min / max, then its lowest chunk and lane by ballots.
// Every lane of a row ends with the row's result.  Returns false, leaving v
// and i as they were, when a candidate of the active lanes is NaN or the row
// has no candidate (the caller then runs row_arg_best).
template <bool Max>
__device__ __forceinline__ bool row_arg_best_lanes(float& v, uint32_t& i, uint32_t chunks,
                                                   uint32_t row_base)
{
    const bool has = i != kNoIndex;
    if (__ballot(has && v != v))
        return false;
    const float none = Max ? -__builtin_inff() : __builtin_inff();
    float x = has ? v : none;
    auto pick = [](float a, float b) { return Max ? fmaxf(a, b) : fminf(a, b); };
    x = pick(x, dpp_f32<0xB1>(x));
    x = pick(x, dpp_f32<0x4E>(x));
    x = pick(x, dpp_f32<0x141>(x));
    x = pick(x, dpp_f32<0x140>(x));
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t at =
            uint32_t(__ballot(has && v == x && (i >> 4) == c) >> row_base) & 0xFFFFu;
        if (at) {
            i = c * 16 + (uint32_t)__ffs(at) - 1u;
            v = x;
            return true;
        }
    }
    return false;
}

// Lanes whose `key` (below 2^Bits) equals this lane's, among the lanes of
// `active`: one ballot per key bit.
template <uint32_t Bits>
__device__ __forceinline__ uint64_t match_lanes(uint32_t key, uint64_t active)
{
    uint64_t same = active;
#pragma unroll
    for (uint32_t b = 0; b < Bits; ++b) {
        const bool set = (key >> b) & 1u;
        const uint64_t m = __ballot(set);
        same &= set ? m : ~m;
    }
    return same;
}

// Lanes of `active` whose `theirs` equals this lane's `mine` (both below
// 2^Bits): one ballot per key bit.
template <uint32_t Bits>
__device__ __forceinline__ uint64_t match_other(uint32_t mine, uint32_t theirs, uint64_t active)
{
    uint64_t same = active;
#pragma unroll
    for (uint32_t b = 0; b < Bits; ++b) {
        const uint64_t m = __ballot((theirs >> b) & 1u);
        same &= ((mine >> b) & 1u) ? m : ~m;
    }
    return same;
}

}  // namespace pg
