// AABB refresh and per-step reductions.
//
// k_compute_aabbs follows Aabb::fromCollider (src/collision/aabb.cpp:8-32):
//   sphere  pos -/+ (r, r, r)
//   box     pos -/+ |R| h with R = toMat3(q), row sums in the reference order
//   hull    component-wise min/max of pos + rotate(q, p_k) over the points
//           (no hull in the reference; this is the builder's definition)
// One thread per collider; the seven pose planes, type and shape planes are
// read coalesced and the six AABB planes written coalesced.
//
// k_reduce reads the AABB planes once and produces what the broad phases
// need: largest x side (SAP row window bound), largest side and mean side
// (grid cell), and centre bounds (LBVH Morton quantisation).  Wave-level
// reductions, one atomic per wave.
#include "devmath.hpp"
#include "internal.hpp"

namespace pg {
namespace {

__global__ void k_step_begin(Counters* c)
{
    uint32_t* w = reinterpret_cast<uint32_t*>(c);
    for (uint32_t i = threadIdx.x; i < sizeof(Counters) / 4; i += blockDim.x)
        w[i] = 0;
    __syncthreads();
    if (threadIdx.x < 3) {
        c->cmin_key[threadIdx.x] = 0xFFFFFFFFu;
        c->cmax_key[threadIdx.x] = 0u;
    }
}

__global__ __launch_bounds__(kBlock) void k_compute_aabbs(
    uint32_t n, const float* __restrict__ pose, const uint8_t* __restrict__ type,
    const float* __restrict__ sdata, const uint32_t* __restrict__ hrange,
    const float* __restrict__ hpts, uint32_t nh, float* __restrict__ aabb)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    pg::pose t{{pose[i], pose[n + i], pose[2 * n + i]},
               {pose[3 * n + i], pose[4 * n + i], pose[5 * n + i], pose[6 * n + i]}};
    const uint8_t ty = type[i];
    f3 mn, mx;
    if (ty == PHYSGPU_SPHERE) {
        const float r = sdata[i];
        f3 rr{r, r, r};
        mn = sub(t.p, rr);
        mx = add(t.p, rr);
    } else if (ty == PHYSGPU_BOX) {
        const f3 h{sdata[i], sdata[n + i], sdata[2 * n + i]};
        const m3 m = to_mat(t.q);
        f3 e{fabsf(m.m00) * h.x + fabsf(m.m01) * h.y + fabsf(m.m02) * h.z,
             fabsf(m.m10) * h.x + fabsf(m.m11) * h.y + fabsf(m.m12) * h.z,
             fabsf(m.m20) * h.x + fabsf(m.m21) * h.y + fabsf(m.m22) * h.z};
        mn = sub(t.p, e);
        mx = add(t.p, e);
    } else {
        const uint32_t first = hrange[i], count = hrange[n + i];
        f3 w = to_world(f3{hpts[first], hpts[nh + first], hpts[2 * nh + first]}, t);
        mn = w;
        mx = w;
        for (uint32_t k = 1; k < count; ++k) {
            const uint32_t p = first + k;
            w = to_world(f3{hpts[p], hpts[nh + p], hpts[2 * nh + p]}, t);
            mn = {min_ref(mn.x, w.x), min_ref(mn.y, w.y), min_ref(mn.z, w.z)};
            mx = {max_ref(mx.x, w.x), max_ref(mx.y, w.y), max_ref(mx.z, w.z)};
        }
    }
    aabb[i] = mn.x;
    aabb[n + i] = mn.y;
    aabb[2 * n + i] = mn.z;
    aabb[3 * n + i] = mx.x;
    aabb[4 * n + i] = mx.y;
    aabb[5 * n + i] = mx.z;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
    for (uint32_t off = 32; off > 0; off >>= 1)
        v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
    for (uint32_t off = 32; off > 0; off >>= 1)
        v = min(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
    for (uint32_t off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    return v;
}

__global__ __launch_bounds__(kBlock) void k_reduce(uint32_t n, const float* __restrict__ aabb,
                                                   Counters* c)
{
    uint32_t wx = 0, wall = 0;
    uint32_t cmn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, cmx[3] = {0u, 0u, 0u};
    unsigned long long wsum = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        float w[3];
        for (int a = 0; a < 3; ++a) {
            const float lo = aabb[a * n + i], hi = aabb[(3 + a) * n + i];
            w[a] = hi - lo;
            // Centre as lo/2 + hi/2 cannot overflow for finite bounds.
            const uint32_t ck = float_key(lo * 0.5f + hi * 0.5f);
            cmn[a] = min(cmn[a], ck);
            cmx[a] = max(cmx[a], ck);
        }
        const float wm = fmaxf(w[0], fmaxf(w[1], w[2]));
        wx = max(wx, float_key(w[0]));
        wall = max(wall, float_key(wm));
        // Sum of sides in 2^-16 m fixed point, clamped, for the mean side.
        const float fx = fminf(wm, 65536.0f) * 65536.0f;
        wsum += (unsigned long long)fx;
    }
    wx = wave_max_u32(wx);
    wall = wave_max_u32(wall);
    wsum = wave_sum_u64(wsum);
    for (int a = 0; a < 3; ++a) {
        cmn[a] = wave_min_u32(cmn[a]);
        cmx[a] = wave_max_u32(cmx[a]);
    }
    if ((threadIdx.x & 63u) == 0) {
        atomicMax(&c->wmax_x_key, wx);
        atomicMax(&c->wmax_key, wall);
        atomicAdd(&c->width_sum_fx, wsum);
        for (int a = 0; a < 3; ++a) {
            atomicMin(&c->cmin_key[a], cmn[a]);
            atomicMax(&c->cmax_key[a], cmx[a]);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_aabb_soa_to_aos(uint32_t n, const float* aabb,
                                                            physgpu_aabb* out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    physgpu_aabb b;
    for (int a = 0; a < 3; ++a) {
        b.min[a] = aabb[a * n + i];
        b.max[a] = aabb[(3 + a) * n + i];
    }
    out[i] = b;
}

__global__ __launch_bounds__(kBlock) void k_aabb_aos_to_soa(uint32_t n, const physgpu_aabb* in,
                                                            float* aabb)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const physgpu_aabb b = in[i];
    for (int a = 0; a < 3; ++a) {
        aabb[a * n + i] = b.min[a];
        aabb[(3 + a) * n + i] = b.max[a];
    }
}

// World pose of each collider: body pose composed with the collider's local
// transform, as physics_world.cpp:660-667 and aabb.cpp:10-13 compose it.
__global__ __launch_bounds__(kBlock) void k_compose(uint32_t n, const float* __restrict__ body,
                                                    const float* __restrict__ local,
                                                    float* __restrict__ pose_out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const f3 bp{body[i], body[n + i], body[2 * n + i]};
    const quat bq{body[3 * n + i], body[4 * n + i], body[5 * n + i], body[6 * n + i]};
    f3 p = bp;
    quat q = bq;
    if (local) {
        const f3 lp{local[i], local[n + i], local[2 * n + i]};
        const quat lq{local[3 * n + i], local[4 * n + i], local[5 * n + i], local[6 * n + i]};
        p = add(bp, rotate(bq, lp));
        q = qmul(bq, lq);
    }
    pose_out[i] = p.x;
    pose_out[n + i] = p.y;
    pose_out[2 * n + i] = p.z;
    pose_out[3 * n + i] = q.x;
    pose_out[4 * n + i] = q.y;
    pose_out[5 * n + i] = q.z;
    pose_out[6 * n + i] = q.w;
}

}  // namespace

void launch_compose_world(physgpu_context* ctx)
{
    if (ctx->n == 0)
        return;
    Timer t;
    time_begin(ctx, "compose_world", t);
    hipLaunchKernelGGL(k_compose, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->n, ctx->world_body_pose,
                       ctx->world_has_local ? ctx->local_pose : nullptr, ctx->pose);
    time_end(ctx, t);
}

void launch_step_begin(physgpu_context* ctx)
{
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, ctx->stream, ctx->counters);
}

void launch_compute_aabbs(physgpu_context* ctx)
{
    if (ctx->n == 0)
        return;
    Timer t;
    time_begin(ctx, "aabb_from_pose", t);
    hipLaunchKernelGGL(k_compute_aabbs, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0,
                       ctx->stream, ctx->n, ctx->pose, ctx->shape_type, ctx->shape_data,
                       ctx->hull_range, ctx->hull_points, ctx->n_hull_points, ctx->aabb);
    time_end(ctx, t);
}

void launch_aabb_reductions(physgpu_context* ctx)
{
    if (ctx->n == 0)
        return;
    Timer t;
    time_begin(ctx, "aabb_reduce", t);
    hipLaunchKernelGGL(k_reduce, dim3(stride_grid(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n,
                       ctx->aabb, ctx->counters);
    time_end(ctx, t);
}

void launch_convert_aabbs_out(physgpu_context* ctx, physgpu_aabb* dev_out)
{
    if (ctx->n == 0)
        return;
    hipLaunchKernelGGL(k_aabb_soa_to_aos, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0,
                       ctx->stream, ctx->n, ctx->aabb, dev_out);
}

void launch_convert_aabbs_in(physgpu_context* ctx, const physgpu_aabb* dev_in)
{
    if (ctx->n == 0)
        return;
    hipLaunchKernelGGL(k_aabb_aos_to_soa, dim3(div_up(ctx->n, kBlock)), dim3(kBlock), 0,
                       ctx->stream, ctx->n, dev_in, ctx->aabb);
}

}  // namespace pg
