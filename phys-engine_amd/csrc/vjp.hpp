// Vector-Jacobian products of the contact geometry's building blocks
// (quaternion rotation, Quaternion::toMat3 columns, normalisation, projected
// radii), shared by the adjoint kernels.
#pragma once

#include "devmath.hpp"

namespace pg {
namespace {

struct PoseGrad {
    f3 p;
    quat q;
};

__device__ __forceinline__ void acc(f3& a, f3 b) { a = add(a, b); }
__device__ __forceinline__ float sgn(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// r = rotate(q, v) = (w^2 - u.u) v + 2 (u.v) u + 2 w (u x v), q = (u, w):
// adds d(gr . r)/dq to gq and returns d(gr . r)/dv = rotate(conj(q), gr).
__device__ f3 rotate_vjp(quat q, f3 v, f3 gr, quat& gq)
{
    const f3 u{q.x, q.y, q.z};
    const float w = q.w;
    const float grv = dot(gr, v), ugr = dot(u, gr), uv = dot(u, v);
    gq.w += 2.0f * w * grv + 2.0f * dot(gr, cross(u, v));
    const f3 gu = add(add(scale(u, -2.0f * grv), scale(v, 2.0f * ugr)),
                      add(scale(gr, 2.0f * uv), scale(cross(v, gr), 2.0f * w)));
    gq.x += gu.x;
    gq.y += gu.y;
    gq.z += gu.z;
    const float s = w * w - dot(u, u);
    return add(add(scale(gr, s), scale(u, 2.0f * ugr)), scale(cross(gr, u), 2.0f * w));
}

// r = rotate(conj(q), v).
__device__ f3 rotate_conj_vjp(quat q, f3 v, f3 gr, quat& gq)
{
    quat gc{0.0f, 0.0f, 0.0f, 0.0f};
    const f3 gv = rotate_vjp(conj(q), v, gr, gc);
    gq.x -= gc.x;
    gq.y -= gc.y;
    gq.z -= gc.z;
    gq.w += gc.w;
    return gv;
}

// Columns of Quaternion::toMat3 (quaternion.hpp:54-64) with gradients g[c].
__device__ void to_mat_vjp(quat q, const f3 g[3], quat& gq)
{
    const float x = q.x, y = q.y, z = q.z, w = q.w;
    const f3 g0 = g[0], g1 = g[1], g2 = g[2];
    gq.x += 2.0f * y * g0.y + 2.0f * z * g0.z + 2.0f * y * g1.x - 4.0f * x * g1.y +
            2.0f * w * g1.z + 2.0f * z * g2.x - 2.0f * w * g2.y - 4.0f * x * g2.z;
    gq.y += -4.0f * y * g0.x + 2.0f * x * g0.y - 2.0f * w * g0.z + 2.0f * x * g1.x +
            2.0f * z * g1.z + 2.0f * w * g2.x + 2.0f * z * g2.y - 4.0f * y * g2.z;
    gq.z += -4.0f * z * g0.x + 2.0f * w * g0.y + 2.0f * x * g0.z - 2.0f * w * g1.x -
            4.0f * z * g1.y + 2.0f * y * g1.z + 2.0f * x * g2.x + 2.0f * y * g2.y;
    gq.w += 2.0f * z * g0.y - 2.0f * y * g0.z - 2.0f * z * g1.x + 2.0f * x * g1.z +
            2.0f * y * g2.x - 2.0f * x * g2.y;
}

// u = a / |a| (len = |a|): returns d/da.
__device__ __forceinline__ f3 unit_vjp(f3 u, float len, f3 gu)
{
    return scale(sub(gu, scale(u, dot(u, gu))), 1.0f / len);
}

// r = sum_k |e_k . axis| h_k  (proj_radius): gradients into axis and e[k].
__device__ void proj_radius_vjp(f3 h, const f3 e[3], f3 axis, float gr, f3& g_axis, f3 g_e[3])
{
    const float hk[3] = {h.x, h.y, h.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c = gr * sgn(dot(e[k], axis)) * hk[k];
        acc(g_axis, scale(e[k], c));
        acc(g_e[k], scale(axis, c));
    }
}

__device__ __forceinline__ void project_tangent(quat& g, quat q)
{
    const float qq = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    const float r = (g.x * q.x + g.y * q.y + g.z * q.z + g.w * q.w) / qq;
    g = {g.x - r * q.x, g.y - r * q.y, g.z - r * q.z, g.w - r * q.w};
}

}  // namespace
}  // namespace pg
