// Adjoint of the narrow phase: gradients of contact normals, penetration
// depths and world contact points with respect to collider positions and
// quaternions, for sphere-sphere, sphere-box / box-sphere and box-box
// manifolds (forward: narrowphase.cpp:132-324, sat.cpp:38-102).
//
// One thread per manifold of the last step.  The forward is recomputed with
// the same FP32 operations as the narrow-phase kernels (contact.hpp), which
// reproduces every discrete choice of the forward — SAT axis and its sign,
// sphere-box branch and nearest face, clamped coordinates, support faces and
// vertices, kept and clamped clipping candidates, the fallback point and its
// overlap-interval ends — and then differentiated in reverse mode through
// the continuous operations with those choices held fixed.
//
// Quaternion gradients are projected onto the tangent space of the unit
// sphere: they are the gradients of the contact as a function of the
// orientation q/|q|.  (Quaternion::toMat3 is only orthogonal for unit q, so
// the radial derivative of |col_k . axis| terms that vanish for unit q is a
// kink with no meaningful value.)
#include "contact.hpp"
#include "devmath.hpp"
#include "internal.hpp"
#include "vjp.hpp"

namespace pg {
namespace {

// Upstream gradients of one manifold and its forward outputs.
struct Seeds {
    f3 normal;
    float depth[4];
    f3 point[4];
};

struct Geometry {
    f3 normal;
    uint32_t count;
    float depth[4];
    f3 point[4];
};

// ------------------------------------------------------------ sphere-sphere
__device__ void sphere_sphere_adjoint(float ra, const pose& ta, float rb, const pose& tb,
                                      const Seeds& g, Geometry& out, PoseGrad& ga, PoseGrad& gb)
{
    const f3 delta = sub(tb.p, ta.p);
    const float dist = length(delta);
    const float radii = ra + rb;
    const bool unit = dist > 1e-6f;
    const f3 n = unit ? divide(delta, dist) : f3{1.0f, 0.0f, 0.0f};
    const f3 pa = add(ta.p, scale(n, ra));
    const f3 pb = sub(tb.p, scale(n, rb));
    out.normal = n;
    out.count = 1;
    out.depth[0] = radii - dist;
    out.point[0] = scale(add(pa, pb), 0.5f);
    // w = (pa + pb) / 2 = (ta.p + tb.p) / 2 + n (ra - rb) / 2
    const f3 gw = scale(g.point[0], 0.5f);
    acc(ga.p, gw);
    acc(gb.p, gw);
    const f3 gn = add(g.normal, scale(g.point[0], 0.5f * (ra - rb)));
    f3 gdelta{0.0f, 0.0f, 0.0f};
    if (unit)
        gdelta = unit_vjp(n, dist, gn);
    if (dist > 0.0f)
        acc(gdelta, scale(divide(delta, dist), -g.depth[0]));
    acc(gb.p, gdelta);
    ga.p = sub(ga.p, gdelta);
}

// ------------------------------------------------------------ sphere-box
__device__ __forceinline__ bool clamp_free(float v, float lo, float hi)
{
    return !(lo == hi) && !(v < lo) && !(v > hi);
}

// Sphere (ts, r) against box (tb, h); the normal points from the sphere to the box.
__device__ void sphere_box_adjoint(float r, const pose& ts, f3 h, const pose& tb, const Seeds& g,
                                   Geometry& out, PoseGrad& gs, PoseGrad& gbx)
{
    const f3 d = sub(ts.p, tb.p);
    const f3 cl = rotate(conj(tb.q), d);
    const f3 closest{clamp_ref(cl.x, -h.x, h.x), clamp_ref(cl.y, -h.y, h.y),
                     clamp_ref(cl.z, -h.z, h.z)};
    const bool fx = clamp_free(cl.x, -h.x, h.x), fy = clamp_free(cl.y, -h.y, h.y),
               fz = clamp_free(cl.z, -h.z, h.z);
    const f3 dl = sub(closest, cl);
    const float dist = length(dl);
    const bool outside = dist > 1e-6f;
    f3 nl;
    float pen;
    int face = 0;
    if (outside) {
        nl = divide(dl, dist);
        pen = r - dist;
    } else {
        // Nearest face, strict <, order +X -X +Y -Y +Z -Z (narrowphase.cpp:198-233).
        const float fv[6] = {h.x - cl.x, h.x + cl.x, h.y - cl.y, h.y + cl.y, h.z - cl.z, h.z + cl.z};
        float nearest = fv[0];
        for (int k = 1; k < 6; ++k)
            if (fv[k] < nearest) {
                nearest = fv[k];
                face = k;
            }
        const float s = (face & 1) ? 1.0f : -1.0f;
        nl = {0.0f, 0.0f, 0.0f};
        if (face < 2)
            nl.x = s;
        else if (face < 4)
            nl.y = s;
        else
            nl.z = s;
        pen = r + nearest;
    }
    const f3 n = rotate(tb.q, nl);
    const f3 box_point = to_world(closest, tb);
    const f3 sphere_point = add(ts.p, scale(n, r));
    out.normal = n;
    out.count = 1;
    out.depth[0] = pen;
    out.point[0] = scale(add(sphere_point, box_point), 0.5f);

    const f3 gsp = scale(g.point[0], 0.5f), gbp = scale(g.point[0], 0.5f);
    acc(gs.p, gsp);
    const f3 gn = add(g.normal, scale(gsp, r));
    acc(gbx.p, gbp);
    f3 gclosest = rotate_vjp(tb.q, closest, gbp, gbx.q);
    const f3 gnl = rotate_vjp(tb.q, nl, gn, gbx.q);
    f3 gcl{0.0f, 0.0f, 0.0f};
    if (outside) {
        const f3 gdl = add(unit_vjp(nl, dist, gnl), scale(nl, -g.depth[0]));
        acc(gclosest, gdl);
        gcl = sub(gcl, gdl);
    } else {
        // pen = r + (h_a -/+ cl_a): the nearest face's coordinate.
        const float s = (face & 1) ? g.depth[0] : -g.depth[0];
        if (face < 2)
            gcl.x += s;
        else if (face < 4)
            gcl.y += s;
        else
            gcl.z += s;
    }
    if (fx)
        gcl.x += gclosest.x;
    if (fy)
        gcl.y += gclosest.y;
    if (fz)
        gcl.z += gclosest.z;
    const f3 gd = rotate_conj_vjp(tb.q, d, gcl, gbx.q);
    acc(gs.p, gd);
    gbx.p = sub(gbx.p, gd);
}

// ------------------------------------------------------------ box-box
// Local corner of supportFaceVertices (narrowphase.cpp:60-98), vertex k.
__device__ f3 face_corner(f3 h, const pose& t, f3 dir, int k)
{
    const f3 l = rotate(conj(t.q), dir);
    int face = 0;
    float largest = fabsf(l.x);
    if (fabsf(l.y) > largest) {
        face = 1;
        largest = fabsf(l.y);
    }
    if (fabsf(l.z) > largest)
        face = 2;
    const float comp = face == 0 ? l.x : (face == 1 ? l.y : l.z);
    const float sign = comp >= 0.0f ? 1.0f : -1.0f;
    const float ext[3] = {h.x, h.y, h.z};
    const int u = (face + 1) % 3, v = (face + 2) % 3;
    float c[3] = {0.0f, 0.0f, 0.0f};
    c[face] = sign * ext[face];
    c[u] = (k & 1) ? ext[u] : -ext[u];
    c[v] = (k & 2) ? ext[v] : -ext[v];
    return {c[0], c[1], c[2]};
}

// Local corner of boxSupportPoint (narrowphase.cpp:27-35).
__device__ __forceinline__ f3 support_corner(f3 h, const pose& t, f3 dir)
{
    const f3 l = rotate(conj(t.q), dir);
    return {l.x >= 0.0f ? h.x : -h.x, l.y >= 0.0f ? h.y : -h.y, l.z >= 0.0f ? h.z : -h.z};
}

struct BoxPair {
    f3 ha, hb;
    pose ta, tb;
    f3 ea[3], eb[3];
};

// overlapMidpoint along t (narrowphase.cpp:48-58): gradients into t, the
// positions and the box axes, with the interval ends the forward chose.
__device__ void overlap_mid_vjp(const BoxPair& P, f3 t, float gom, f3& gt, PoseGrad& ga,
                                PoseGrad& gb, f3 gea[3], f3 geb[3])
{
    const float ca = dot(P.ta.p, t), cb = dot(P.tb.p, t);
    const float hka[3] = {P.ha.x, P.ha.y, P.ha.z}, hkb[3] = {P.hb.x, P.hb.y, P.hb.z};
    const float ra = fabsf(dot(P.ea[0], t)) * hka[0] + fabsf(dot(P.ea[1], t)) * hka[1] +
                     fabsf(dot(P.ea[2], t)) * hka[2];
    const float rb = fabsf(dot(P.eb[0], t)) * hkb[0] + fabsf(dot(P.eb[1], t)) * hkb[1] +
                     fabsf(dot(P.eb[2], t)) * hkb[2];
    const float half = 0.5f * gom;
    float gca = 0.0f, gcb = 0.0f, gra = 0.0f, grb = 0.0f;
    // lower = max(ca - ra, cb - rb): the second wins only when strictly larger.
    if ((ca - ra) < (cb - rb)) {
        gcb += half;
        grb -= half;
    } else {
        gca += half;
        gra -= half;
    }
    // upper = min(ca + ra, cb + rb): the second wins only when strictly smaller.
    if ((cb + rb) < (ca + ra)) {
        gcb += half;
        grb += half;
    } else {
        gca += half;
        gra += half;
    }
    acc(ga.p, scale(t, gca));
    acc(gt, scale(P.ta.p, gca));
    acc(gb.p, scale(t, gcb));
    acc(gt, scale(P.tb.p, gcb));
    proj_radius_vjp(P.ha, P.ea, t, gra, gt, gea);
    proj_radius_vjp(P.hb, P.eb, t, grb, gt, geb);
}

__device__ void box_box_adjoint(f3 ha, const pose& ta, f3 hb, const pose& tb, const BoxTape& tp,
                                const Seeds& g, Geometry& out, PoseGrad& ga, PoseGrad& gb)
{
    BoxPair P;
    P.ha = ha;
    P.hb = hb;
    P.ta = ta;
    P.tb = tb;
    const m3 ma = to_mat(ta.q), mb = to_mat(tb.q);
    P.ea[0] = col0(ma);
    P.ea[1] = col1(ma);
    P.ea[2] = col2(ma);
    P.eb[0] = col0(mb);
    P.eb[1] = col1(mb);
    P.eb[2] = col2(mb);
    const f3 dab = sub(tb.p, ta.p);
    const uint32_t sel = tp.sel;
    const uint32_t ci = sel >= 6 ? (sel - 6) / 3 : 0, cj = sel >= 6 ? (sel - 6) % 3 : 0;
    const f3 axis = sel < 3 ? P.ea[sel] : (sel < 6 ? P.eb[sel - 3] : cross(P.ea[ci], P.eb[cj]));
    const float lsq = dot(axis, axis);
    const float len = sqrtf(lsq);
    const f3 unit = scale(axis, 1.0f / len);
    const f3 n = tp.flip ? neg(unit) : unit;
    out.normal = n;
    out.count = tp.count;

    f3 gea[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, geb[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    f3 gn = g.normal;
    float gpen = 0.0f;
    if (tp.fallback) {
        // depth = SAT depth; point = n nc + t1 om(t1) + t2 om(t2) (narrowphase.cpp:270-276).
        const float pen = (proj_radius(ha, ma, unit) + proj_radius(hb, mb, unit)) -
                          fabsf(dot(dab, unit));
        out.depth[0] = pen;
        gpen += g.depth[0];
        const f3 ref = fabsf(n.x) < 0.9f ? f3{1.0f, 0.0f, 0.0f} : f3{0.0f, 1.0f, 0.0f};
        const f3 c1 = cross(n, ref);
        const float len1 = length(c1);
        const bool t1_ok = !(len1 <= 1e-8f);
        const f3 t1 = t1_ok ? scale(c1, 1.0f / len1) : f3{0.0f, 0.0f, 0.0f};
        const f3 t2 = cross(n, t1);
        const float nc = (dot(ta.p, n) + proj_radius(ha, ma, n) + dot(tb.p, n) -
                          proj_radius(hb, mb, n)) * 0.5f;
        const float om1 = overlap_mid(ha, ta, ma, hb, tb, mb, t1);
        const float om2 = overlap_mid(ha, ta, ma, hb, tb, mb, t2);
        out.point[0] = add(add(scale(n, nc), scale(t1, om1)), scale(t2, om2));
        const f3 gfb = g.point[0];
        acc(gn, scale(gfb, nc));
        const float gnc = dot(gfb, n);
        f3 gt1 = scale(gfb, om1), gt2 = scale(gfb, om2);
        overlap_mid_vjp(P, t1, dot(gfb, t1), gt1, ga, gb, gea, geb);
        overlap_mid_vjp(P, t2, dot(gfb, t2), gt2, ga, gb, gea, geb);
        // t2 = n x t1
        acc(gn, cross(t1, gt2));
        acc(gt1, cross(gt2, n));
        if (t1_ok) {
            const f3 gc1 = unit_vjp(t1, len1, gt1);
            acc(gn, cross(ref, gc1));  // c1 = n x ref
        }
        const float hn = 0.5f * gnc;
        acc(ga.p, scale(n, hn));
        acc(gn, scale(ta.p, hn));
        acc(gb.p, scale(n, hn));
        acc(gn, scale(tb.p, hn));
        proj_radius_vjp(ha, P.ea, n, hn, gn, gea);
        proj_radius_vjp(hb, P.eb, n, -hn, gn, geb);
    } else {
        for (uint32_t k = 0; k < tp.count; ++k) {
            const bool from_a = tp.side[k] == 0;
            const f3 dir = from_a ? n : neg(n);
            const f3 hs = from_a ? ha : hb, ho = from_a ? hb : ha;
            const pose& ts = from_a ? ta : tb;
            const pose& to = from_a ? tb : ta;
            PoseGrad& gsrc = from_a ? ga : gb;
            PoseGrad& goth = from_a ? gb : ga;
            const f3 c = face_corner(hs, ts, dir, tp.corner[k]);
            const f3 v = to_world(c, ts);
            const f3 sv = support_corner(ho, to, neg(dir));
            const f3 sp = to_world(sv, to);
            const float os = dot(sp, dir);
            const float excess = dot(v, dir) - os;
            const float d = tp.zero[k] ? 0.0f : excess;
            out.depth[k] = d;
            out.point[k] = sub(v, scale(dir, d * 0.5f));
            // cand = v - dir d / 2
            const f3 gc = g.point[k];
            f3 gv = gc;
            f3 gdir = scale(gc, -0.5f * d);
            float gd = g.depth[k] - 0.5f * dot(gc, dir);
            if (!tp.zero[k]) {
                // d = v . dir - sp . dir
                acc(gv, scale(dir, gd));
                acc(gdir, scale(v, gd));
                const f3 gsp = scale(dir, -gd);
                acc(gdir, scale(sp, -gd));
                acc(goth.p, gsp);
                rotate_vjp(to.q, sv, gsp, goth.q);
            }
            acc(gsrc.p, gv);
            rotate_vjp(ts.q, c, gv, gsrc.q);
            acc(gn, from_a ? gdir : neg(gdir));
        }
    }
    // n = +/- unit; SAT depth = ra(unit) + rb(unit) - |dab . unit|.
    f3 gunit = tp.flip ? neg(gn) : gn;
    if (gpen != 0.0f) {
        proj_radius_vjp(ha, P.ea, unit, gpen, gunit, gea);
        proj_radius_vjp(hb, P.eb, unit, gpen, gunit, geb);
        const float s = sgn(dot(dab, unit));
        acc(gunit, scale(dab, -gpen * s));
        const f3 gdab = scale(unit, -gpen * s);
        acc(gb.p, gdab);
        ga.p = sub(ga.p, gdab);
    }
    // unit = axis / |axis|; the axis is a column of A or B or A_i x B_j.
    const f3 gaxis = unit_vjp(unit, len, gunit);
    if (sel < 3) {
        acc(gea[sel], gaxis);
    } else if (sel < 6) {
        acc(geb[sel - 3], gaxis);
    } else {
        acc(gea[ci], cross(P.eb[cj], gaxis));
        acc(geb[cj], cross(gaxis, P.ea[ci]));
    }
    to_mat_vjp(ta.q, gea, ga.q);
    to_mat_vjp(tb.q, geb, gb.q);
}

// Gradient planes: pose 7 x n (px py pz qx qy qz qw); pair 14 x m (A's 7, B's 7).
__global__ __launch_bounds__(kBlock) void k_contact_adjoint(
    uint32_t n, const float4* __restrict__ pose4, const float4* __restrict__ shape4,
    const physgpu_manifold* __restrict__ man, uint32_t m,
    const Counters* __restrict__ counters, const float* __restrict__ g_normal,
    const float* __restrict__ g_depth, const float* __restrict__ g_point,
    float* __restrict__ geom_normal, float* __restrict__ geom_depth,
    float* __restrict__ geom_point, float* __restrict__ grad_pair, float* __restrict__ grad_pose,
    uint32_t* __restrict__ skipped)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = min(m, counters->manifold_count);
    if (t >= m)
        return;
    Geometry out{};
    PoseGrad ga{{0, 0, 0}, {0, 0, 0, 0}}, gb{{0, 0, 0}, {0, 0, 0, 0}};
    bool done = false;
    uint32_t a = 0, b = 0;
    if (t < count) {
        const physgpu_manifold& mf = man[t];
        a = mf.first;
        b = mf.second;
        const uint32_t cls = mf.shape_pair;
        Seeds g{};
        if (g_normal)
            g.normal = {g_normal[t], g_normal[m + t], g_normal[2 * m + t]};
        for (int k = 0; k < 4; ++k) {
            g.depth[k] = g_depth ? g_depth[size_t(k) * m + t] : 0.0f;
            g.point[k] = g_point ? f3{g_point[size_t(3 * k) * m + t],
                                      g_point[size_t(3 * k + 1) * m + t],
                                      g_point[size_t(3 * k + 2) * m + t]}
                                 : f3{0.0f, 0.0f, 0.0f};
        }
        for (int k = mf.point_count; k < 4; ++k) {
            g.depth[k] = 0.0f;
            g.point[k] = {0.0f, 0.0f, 0.0f};
        }
        auto data = [&](uint32_t i) { return shape_data(shape4, i); };
        const pose ta = load_pose4(pose4, a), tb = load_pose4(pose4, b);
        if (cls == PHYSGPU_PAIR_SPHERE_SPHERE) {
            sphere_sphere_adjoint(data(a).x, ta, data(b).x, tb, g, out, ga, gb);
            done = true;
        } else if (cls == PHYSGPU_PAIR_SPHERE_BOX) {
            if (shape_type(shape4, a) == PHYSGPU_SPHERE) {
                sphere_box_adjoint(data(a).x, ta, data(b), tb, g, out, ga, gb);
            } else {
                // Box-sphere: the forward swaps and negates the normal (:353-365).
                Seeds gs = g;
                gs.normal = neg(g.normal);
                sphere_box_adjoint(data(b).x, tb, data(a), ta, gs, out, gb, ga);
                out.normal = neg(out.normal);
            }
            done = true;
        } else if (cls == PHYSGPU_PAIR_BOX_BOX) {
            BoxTape tp{};
            Contact ct;
            const f3 ha = data(a), hb = data(b);
            if (box_box(ha, ta, hb, tb, ct, &tp)) {
                box_box_adjoint(ha, ta, hb, tb, tp, g, out, ga, gb);
                done = true;
            }
        }
        if (done) {
            project_tangent(ga.q, ta.q);
            project_tangent(gb.q, tb.q);
        }
    }
    if (geom_normal) {
        geom_normal[t] = out.normal.x;
        geom_normal[m + t] = out.normal.y;
        geom_normal[2 * m + t] = out.normal.z;
    }
    for (int k = 0; k < 4; ++k) {
        const bool used = k < (int)out.count;
        if (geom_depth)
            geom_depth[size_t(k) * m + t] = used ? out.depth[k] : 0.0f;
        if (geom_point) {
            geom_point[size_t(3 * k) * m + t] = used ? out.point[k].x : 0.0f;
            geom_point[size_t(3 * k + 1) * m + t] = used ? out.point[k].y : 0.0f;
            geom_point[size_t(3 * k + 2) * m + t] = used ? out.point[k].z : 0.0f;
        }
    }
    const float va[7] = {ga.p.x, ga.p.y, ga.p.z, ga.q.x, ga.q.y, ga.q.z, ga.q.w};
    const float vb[7] = {gb.p.x, gb.p.y, gb.p.z, gb.q.x, gb.q.y, gb.q.z, gb.q.w};
    if (grad_pair)
        for (int c = 0; c < 7; ++c) {
            grad_pair[size_t(c) * m + t] = va[c];
            grad_pair[size_t(7 + c) * m + t] = vb[c];
        }
    if (grad_pose && done)
        for (int c = 0; c < 7; ++c) {
            atomicAdd(&grad_pose[size_t(c) * n + a], va[c]);
            atomicAdd(&grad_pose[size_t(c) * n + b], vb[c]);
        }
    // Hull manifolds are handled (or counted as skipped) by k_hull_adjoint.
    if (t < count && !done && man[t].shape_pair < PHYSGPU_PAIR_SPHERE_HULL)
        atomicAdd(skipped, 1u);
}

__global__ __launch_bounds__(kBlock) void k_zero(float* __restrict__ p, uint64_t count,
                                                 uint32_t* __restrict__ flag)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *flag = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

}  // namespace

void launch_contact_adjoint(physgpu_context* ctx, uint32_t m, const float* g_normal,
                            const float* g_depth, const float* g_point, float* geom_normal,
                            float* geom_depth, float* geom_point, float* grad_pair,
                            float* grad_pose, uint32_t* skipped)
{
    hipLaunchKernelGGL(k_zero, dim3(stride_grid(grad_pose ? 7ull * ctx->n : 1)), dim3(kBlock), 0,
                       ctx->stream, grad_pose ? grad_pose : geom_normal,
                       grad_pose ? 7ull * ctx->n : 0ull, skipped);
    if (m == 0)
        return;
    Timer t;
    time_begin(ctx, "np_adjoint", t);
    hipLaunchKernelGGL(k_contact_adjoint, dim3(div_up(m, kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->n, ctx->pose4, ctx->shape4, ctx->manifolds, m,
                       ctx->counters, g_normal, g_depth, g_point, geom_normal, geom_depth,
                       geom_point, grad_pair, grad_pose, skipped);
    time_end(ctx, t);
    launch_hull_adjoint(ctx, m, g_normal, g_depth, g_point, geom_normal, geom_depth, geom_point,
                        grad_pair, grad_pose, skipped);
}

}  // namespace pg
