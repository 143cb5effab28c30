// CPU restatement of the phys-engine collision path — TEST INFRASTRUCTURE
// (the oracle and the "port" CPU baseline; never linked into the product).
//
// Reference files followed (paths relative to the reference root):
//   AABB            src/collision/aabb.cpp:8-32
//   SAP             src/collision/broadphase.cpp:195-291 (+ ParallelFor ranges)
//   pair ordering   src/collision/broadphase.cpp:18-60
//   uniform grid    src/collision/broadphase.cpp:62-182
//   tree pairs      src/collision/dynamic_aabb_tree.cpp:268-336 (pair set only:
//                   sorted unique pairs of overlapping tight bounds)
//   narrow phase    src/collision/narrowphase.cpp:16-374, sat.cpp:38-102
//   GJK / EPA       src/collision/gjk.cpp:11-158, epa.cpp:29-130
//   world section   src/world/physics_world.cpp:539-739
//   worker pool     src/world/step_workspace.hpp:24-172
// Hull AABBs, hull/sphere/box support maps and the GJK/EPA contact point are
// defined by this project (the reference has no hull shape and no GJK/EPA
// caller); they are stated once here and once in the HIP kernels.
#include "physcpu.hpp"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>

#include "../include/physgpu.h"

namespace orc {

// Floating-point operation counts of the calling thread (only the Counted
// scalar below adds to them) and GJK / EPA iteration counts (every scalar
// type), for the narrow phase's operation count (orc_count_flops).
struct FlopCount {
    uint64_t add = 0, mul = 0, div = 0, sqrt = 0;
    uint64_t total() const { return add + mul + div + sqrt; }
};
thread_local FlopCount t_flops;
thread_local uint64_t t_gjk_iterations = 0, t_epa_iterations = 0, t_gjk_flops = 0;

// ---------------------------------------------------------------- AABB
template <typename T>
struct Box6 {
    V3<T> mn, mx;
};

// Aabb::fromCollider, aabb.cpp:8-32 (world pose already composed); hulls take
// the component-wise min/max of their points mapped by toWorld.
template <typename T>
Box6<T> aabbOf(const Shape<T>& s, const Pose<T>& t)
{
    if (s.type == kSphere) {
        V3<T> r{s.data.x, s.data.x, s.data.x};
        return {sub(t.p, r), add(t.p, r)};
    }
    if (s.type == kBox) {
        M3<T> m = toMat(t.q);
        V3<T> h = s.data;
        V3<T> e{absOf(m.m[0][0]) * h.x + absOf(m.m[0][1]) * h.y + absOf(m.m[0][2]) * h.z,
                absOf(m.m[1][0]) * h.x + absOf(m.m[1][1]) * h.y + absOf(m.m[1][2]) * h.z,
                absOf(m.m[2][0]) * h.x + absOf(m.m[2][1]) * h.y + absOf(m.m[2][2]) * h.z};
        return {sub(t.p, e), add(t.p, e)};
    }
    V3<T> w = toWorld(s.points[0], t);
    Box6<T> b{w, w};
    for (uint32_t k = 1; k < s.pointCount; ++k) {
        w = toWorld(s.points[k], t);
        b.mn = {minOf(b.mn.x, w.x), minOf(b.mn.y, w.y), minOf(b.mn.z, w.z)};
        b.mx = {maxOf(b.mx.x, w.x), maxOf(b.mx.y, w.y), maxOf(b.mx.z, w.z)};
    }
    return b;
}

// ---------------------------------------------------------------- pool
// A fixed pool with an atomic chunk counter; the caller thread takes part
// (step_workspace.hpp:54-92 semantics: chunks of `chunk` items).
class Pool {
public:
    template <typename F>
    static void run(uint32_t workers, size_t items, F&& f, size_t chunk = 64)
    {
        if (workers <= 1 || items == 0) {
            for (size_t i = 0; i < items; ++i)
                f(i);
            return;
        }
        std::atomic<size_t> next{0};
        std::exception_ptr failure;
        std::mutex failureMutex;
        auto body = [&] {
            try {
                for (;;) {
                    size_t begin = next.fetch_add(chunk);
                    if (begin >= items)
                        break;
                    size_t end = std::min(items, begin + chunk);
                    for (size_t i = begin; i < end; ++i)
                        f(i);
                }
            } catch (...) {
                std::lock_guard<std::mutex> lock(failureMutex);
                if (!failure)
                    failure = std::current_exception();
            }
        };
        std::vector<std::thread> threads;
        for (uint32_t w = 1; w < workers; ++w)
            threads.emplace_back(body);
        body();
        for (auto& t : threads)
            t.join();
        if (failure)
            std::rethrow_exception(failure);
    }
};

// ---------------------------------------------------------------- pairs
struct Pair {
    uint32_t first, second;
};

inline bool overlaps(const physgpu_aabb& a, const physgpu_aabb& b)
{
    return a.min[0] <= b.max[0] && a.max[0] >= b.min[0] && a.min[1] <= b.max[1] &&
           a.max[1] >= b.min[1] && a.min[2] <= b.max[2] && a.max[2] >= b.min[2];
}

inline bool pairLess(const Pair& l, const Pair& r)
{
    return l.first != r.first ? l.first < r.first : l.second < r.second;
}

// sortPairsLinear, broadphase.cpp:27-60: two stable counting passes (by
// second, then by first) when there are at least as many pairs as keys.
void orderPairs(std::vector<Pair>& pairs, size_t keys)
{
    if (pairs.size() < 2)
        return;
    if (pairs.size() < keys) {
        std::sort(pairs.begin(), pairs.end(), pairLess);
        return;
    }
    std::vector<Pair> tmp(pairs.size());
    std::vector<size_t> count(keys);
    auto pass = [&](const std::vector<Pair>& in, std::vector<Pair>& out, bool byFirst) {
        std::fill(count.begin(), count.end(), 0);
        for (const Pair& p : in)
            ++count[byFirst ? p.first : p.second];
        size_t run = 0;
        for (size_t& c : count) {
            size_t f = c;
            c = run;
            run += f;
        }
        for (const Pair& p : in)
            out[count[byFirst ? p.first : p.second]++] = p;
    };
    pass(pairs, tmp, false);
    pass(tmp, pairs, true);
}

struct SweepRecord {
    physgpu_aabb b;
    uint32_t index;
};

// SAP, broadphase.cpp:205-291.
void sweepAndPrune(const physgpu_aabb* bounds, uint32_t n, uint32_t workers,
                   std::vector<Pair>& pairs, uint64_t& windowTests)
{
    std::vector<SweepRecord> rec(n);
    for (uint32_t i = 0; i < n; ++i)
        rec[i] = {bounds[i], i};
    std::sort(rec.begin(), rec.end(), [](const SweepRecord& l, const SweepRecord& r) {
        if (l.b.min[0] != r.b.min[0])
            return l.b.min[0] < r.b.min[0];
        return l.index < r.index;
    });
    auto sweep = [&](size_t begin, size_t end, std::vector<Pair>& out, uint64_t& tests) {
        for (size_t i = begin; i < end; ++i) {
            const SweepRecord& a = rec[i];
            size_t j = i + 1;
            for (; j < rec.size(); ++j) {
                const SweepRecord& b = rec[j];
                if (b.b.min[0] > a.b.max[0])
                    break;
                if (a.b.min[1] <= b.b.max[1] && a.b.max[1] >= b.b.min[1] &&
                    a.b.min[2] <= b.b.max[2] && a.b.max[2] >= b.b.min[2])
                    out.push_back({std::min(a.index, b.index), std::max(a.index, b.index)});
            }
            tests += j - i - 1;
        }
    };
    pairs.clear();
    windowTests = 0;
    if (workers > 1 && n >= 4096) {
        const size_t perWorker = 16;
        size_t ranges = workers > n / perWorker ? n : size_t(workers) * perWorker;
        std::vector<std::vector<Pair>> parts(ranges);
        std::vector<uint64_t> tests(ranges, 0);
        Pool::run(workers, ranges, [&](size_t r) {
            sweep(size_t(n) * r / ranges, size_t(n) * (r + 1) / ranges, parts[r], tests[r]);
        }, 1);
        for (size_t r = 0; r < ranges; ++r) {
            pairs.insert(pairs.end(), parts[r].begin(), parts[r].end());
            windowTests += tests[r];
        }
    } else {
        sweep(0, n, pairs, windowTests);
    }
    orderPairs(pairs, n);
}

// gridRange, broadphase.cpp:62-88.
bool cellRange(const physgpu_aabb& b, double cell, int32_t first[3], int counts[3])
{
    const size_t maxCells = 64;
    size_t total = 1;
    for (int a = 0; a < 3; ++a) {
        double lo = std::floor(b.min[a] / cell);
        double hi = std::floor(b.max[a] / cell);
        if (!std::isfinite(lo) || !std::isfinite(hi) ||
            lo < double(std::numeric_limits<int32_t>::min()) ||
            hi > double(std::numeric_limits<int32_t>::max()) || lo > hi)
            return false;
        int64_t c = int64_t(hi) - int64_t(lo) + 1;
        if (c > int64_t(maxCells / total))
            return false;
        first[a] = int32_t(lo);
        counts[a] = int(c);
        total *= size_t(c);
    }
    return true;
}

struct CellRecord {
    int32_t cell[3];
    uint32_t index;
};

// findGridPairs, broadphase.cpp:90-182, including its statistics.
void uniformGrid(const physgpu_aabb* bounds, uint32_t n, std::vector<Pair>& pairs,
                 physgpu_broadphase_stats& st)
{
    std::vector<double> widths;
    widths.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        double w = std::max({double(bounds[i].max[0]) - bounds[i].min[0],
                             double(bounds[i].max[1]) - bounds[i].min[1],
                             double(bounds[i].max[2]) - bounds[i].min[2]});
        if (w > 0.0 && std::isfinite(w))
            widths.push_back(w);
    }
    double cell = 1.0;
    if (!widths.empty()) {
        auto mid = widths.begin() + widths.size() / 2;
        std::nth_element(widths.begin(), mid, widths.end());
        cell = *mid;
    }
    std::vector<CellRecord> entries;
    std::vector<uint8_t> inGrid(n, 0);
    std::vector<uint32_t> overflow;
    for (uint32_t i = 0; i < n; ++i) {
        int32_t f[3];
        int c[3];
        if (!cellRange(bounds[i], cell, f, c)) {
            overflow.push_back(i);
            continue;
        }
        inGrid[i] = 1;
        for (int x = 0; x < c[0]; ++x)
            for (int y = 0; y < c[1]; ++y)
                for (int z = 0; z < c[2]; ++z)
                    entries.push_back({{f[0] + x, f[1] + y, f[2] + z}, i});
    }
    std::sort(entries.begin(), entries.end(), [](const CellRecord& l, const CellRecord& r) {
        for (int a = 0; a < 3; ++a)
            if (l.cell[a] != r.cell[a])
                return l.cell[a] < r.cell[a];
        return l.index < r.index;
    });
    uint64_t comparisons = 0;
    pairs.clear();
    auto check = [&](uint32_t a, uint32_t b) {
        ++comparisons;
        if (overlaps(bounds[a], bounds[b]))
            pairs.push_back({std::min(a, b), std::max(a, b)});
    };
    for (size_t begin = 0; begin < entries.size();) {
        size_t end = begin + 1;
        while (end < entries.size() && std::memcmp(entries[end].cell, entries[begin].cell,
                                                   sizeof(entries[begin].cell)) == 0)
            ++end;
        for (size_t i = begin; i < end; ++i)
            for (size_t j = i + 1; j < end; ++j)
                check(entries[i].index, entries[j].index);
        begin = end;
    }
    for (uint32_t o : overflow)
        for (uint32_t j = 0; j < n; ++j)
            if (o != j && (inGrid[j] || o < j))
                check(o, j);
    std::sort(pairs.begin(), pairs.end(), pairLess);
    pairs.erase(std::unique(pairs.begin(), pairs.end(),
                            [](const Pair& l, const Pair& r) {
                                return l.first == r.first && l.second == r.second;
                            }),
                pairs.end());
    st.aabb_pairs = pairs.size();
    st.grid_comparisons = comparisons;
    st.grid_entries = entries.size();
    st.grid_overflow_aabbs = overflow.size();
    st.grid_cell_size = cell;
}

// ---------------------------------------------------------------- narrow phase
namespace np {

// boxSupportPoint, narrowphase.cpp:27-35.
template <typename T>
V3<T> boxSupport(V3<T> h, const Pose<T>& t, V3<T> dir)
{
    V3<T> l = rotate(conj(t.q), dir);
    V3<T> p{l.x >= T(0) ? h.x : -h.x, l.y >= T(0) ? h.y : -h.y, l.z >= T(0) ? h.z : -h.z};
    return toWorld(p, t);
}

// boxProjectedRadius, narrowphase.cpp:37-46 (same sum as sat.cpp:38-42).
template <typename T>
T projectedRadius(V3<T> h, const Pose<T>& t, V3<T> axis)
{
    M3<T> m = toMat(t.q);
    return absOf(dot(column(m, 0), axis)) * h.x + absOf(dot(column(m, 1), axis)) * h.y +
           absOf(dot(column(m, 2), axis)) * h.z;
}

// overlapMidpoint, narrowphase.cpp:48-58.
template <typename T>
T overlapMid(V3<T> ha, const Pose<T>& ta, V3<T> hb, const Pose<T>& tb, V3<T> axis)
{
    T ca = dot(ta.p, axis);
    T cb = dot(tb.p, axis);
    T ra = projectedRadius(ha, ta, axis);
    T rb = projectedRadius(hb, tb, axis);
    T lower = maxOf(ca - ra, cb - rb);
    T upper = minOf(ca + ra, cb + rb);
    return (lower + upper) * T(0.5f);
}

// supportFaceVertices, narrowphase.cpp:60-98.
template <typename T>
void supportFace(V3<T> h, const Pose<T>& t, V3<T> dir, V3<T> out[4])
{
    V3<T> l = rotate(conj(t.q), dir);
    int face = 0;
    T largest = absOf(l.x);
    if (absOf(l.y) > largest) {
        face = 1;
        largest = absOf(l.y);
    }
    if (absOf(l.z) > largest)
        face = 2;
    T comp[3] = {l.x, l.y, l.z};
    T sign = comp[face] >= T(0) ? T(1) : T(-1);
    T ext[3] = {h.x, h.y, h.z};
    int u = (face + 1) % 3, v = (face + 2) % 3;
    for (int k = 0; k < 4; ++k) {
        T c[3] = {T(0), T(0), T(0)};
        c[face] = sign * ext[face];
        c[u] = (k & 1) ? ext[u] : -ext[u];
        c[v] = (k & 2) ? ext[v] : -ext[v];
        out[k] = toWorld(V3<T>{c[0], c[1], c[2]}, t);
    }
}

// pointInsideBox, narrowphase.cpp:100-107.
template <typename T>
bool insideBox(V3<T> p, V3<T> h, const Pose<T>& t)
{
    V3<T> l = rotate(conj(t.q), sub(p, t.p));
    const T tol = T(1e-4f);
    return absOf(l.x) <= h.x + tol && absOf(l.y) <= h.y + tol && absOf(l.z) <= h.z + tol;
}

// duplicatePoint, narrowphase.cpp:109-120.
template <typename T>
bool duplicate(const V3<T>* pts, uint32_t count, V3<T> p)
{
    for (uint32_t i = 0; i < count; ++i) {
        V3<T> d = sub(pts[i], p);
        if (dot(d, d) <= T(1e-6f))
            return true;
    }
    return false;
}

// overlapOnAxis, sat.cpp:45-62.
template <typename T>
bool axisOverlap(V3<T> ha, const Pose<T>& ta, V3<T> hb, const Pose<T>& tb, V3<T> axis,
                 V3<T>& unit, T& depth)
{
    T lsq = dot(axis, axis);
    if (lsq < T(1e-8f)) {
        depth = bigOf<T>();
        return true;
    }
    unit = scale(axis, T(1) / sqrtOf(lsq));
    T dist = absOf(dot(sub(tb.p, ta.p), unit));
    T d = (projectedRadius(ha, ta, unit) + projectedRadius(hb, tb, unit)) - dist;
    if (d < T(0))
        return false;
    depth = d;
    return true;
}

// testOBBOBB, sat.cpp:66-102.  The depth compared for the degenerate-axis
// case is FLT_MAX in FP32; the FP64 twin keeps FLT_MAX too (never selected).
template <typename T>
bool sat(V3<T> ha, const Pose<T>& ta, V3<T> hb, const Pose<T>& tb, V3<T>& axisOut, T& depthOut,
         int* selOut = nullptr)
{
    int index = 0;
    depthOut = T(std::numeric_limits<float>::max());
    M3<T> ma = toMat(ta.q), mb = toMat(tb.q);
    V3<T> ea[3] = {column(ma, 0), column(ma, 1), column(ma, 2)};
    V3<T> eb[3] = {column(mb, 0), column(mb, 1), column(mb, 2)};
    auto consider = [&](V3<T> axis) {
        V3<T> unit{};
        T d;
        if (!axisOverlap(ha, ta, hb, tb, axis, unit, d))
            return false;
        if (d < depthOut) {
            depthOut = d;
            axisOut = unit;
            if (selOut)
                *selOut = index;
        }
        ++index;
        return true;
    };
    for (int i = 0; i < 3; ++i)
        if (!consider(ea[i]))
            return false;
    for (int i = 0; i < 3; ++i)
        if (!consider(eb[i]))
            return false;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (!consider(cross(ea[i], eb[j])))
                return false;
    return true;
}

template <typename T>
void setAnchors(Contact<T>& c, uint32_t k, V3<T> world, const Pose<T>& ta, const Pose<T>& tb)
{
    c.anchorA[k] = toLocal(world, ta);
    c.anchorB[k] = toLocal(world, tb);
    c.world[k] = world;
}

// intersectSphereSphere, narrowphase.cpp:132-160.
template <typename T>
bool sphereSphere(T ra, const Pose<T>& ta, T rb, const Pose<T>& tb, Contact<T>& c)
{
    V3<T> delta = sub(tb.p, ta.p);
    T dist = length(delta);
    T radii = ra + rb;
    if (dist >= radii)
        return false;
    c.normal = dist > T(1e-6f) ? divide(delta, dist) : V3<T>{T(1), T(0), T(0)};
    c.count = 1;
    c.depth[0] = radii - dist;
    V3<T> pa = add(ta.p, scale(c.normal, ra));
    V3<T> pb = sub(tb.p, scale(c.normal, rb));
    V3<T> w = scale(add(pa, pb), T(0.5f));
    setAnchors(c, 0, w, ta, tb);
    return true;
}

// Math3d::clamp, math_utils.hpp:31-39 (min > max cannot occur: extents >= 0).
template <typename T>
T clampTo(T v, T lo, T hi)
{
    if (lo == hi)
        return lo;
    if (lo > hi)
        throw std::invalid_argument("clamp minimum exceeds maximum");
    if (v < lo)
        return lo;
    if (v > hi)
        return hi;
    return v;
}

// intersectSphereBox, narrowphase.cpp:162-248.
template <typename T>
bool sphereBox(T r, const Pose<T>& ts, V3<T> h, const Pose<T>& tb, Contact<T>& c,
               Choice* ch = nullptr)
{
    V3<T> cl = rotate(conj(tb.q), sub(ts.p, tb.p));
    V3<T> closest{clampTo(cl.x, -h.x, h.x), clampTo(cl.y, -h.y, h.y), clampTo(cl.z, -h.z, h.z)};
    if (ch)
        ch->clamped = (closest.x != cl.x ? 1 : 0) | (closest.y != cl.y ? 2 : 0) |
                      (closest.z != cl.z ? 4 : 0);
    V3<T> dl = sub(closest, cl);
    T dist = length(dl);
    V3<T> nl;
    T pen;
    if (dist > T(1e-6f)) {
        if (dist >= r)
            return false;
        nl = divide(dl, dist);
        pen = r - dist;
    } else {
        T face[6] = {h.x - cl.x, h.x + cl.x, h.y - cl.y, h.y + cl.y, h.z - cl.z, h.z + cl.z};
        const V3<T> normals[6] = {{T(-1), T(0), T(0)}, {T(1), T(0), T(0)}, {T(0), T(-1), T(0)},
                                  {T(0), T(1), T(0)},  {T(0), T(0), T(-1)}, {T(0), T(0), T(1)}};
        T nearest = face[0];
        nl = normals[0];
        int chosen = 0;
        for (int k = 1; k < 6; ++k)
            if (face[k] < nearest) {
                nearest = face[k];
                nl = normals[k];
                chosen = k;
            }
        pen = r + nearest;
        if (ch) {
            ch->branch = 1;
            ch->face = chosen;
        }
    }
    c.normal = rotate(tb.q, nl);
    c.count = 1;
    c.depth[0] = pen;
    V3<T> boxPoint = toWorld(closest, tb);
    V3<T> spherePoint = add(ts.p, scale(c.normal, r));
    V3<T> w = scale(add(spherePoint, boxPoint), T(0.5f));
    setAnchors(c, 0, w, ts, tb);
    return true;
}

// intersectBoxBox, narrowphase.cpp:250-324.
template <typename T>
bool boxBox(V3<T> ha, const Pose<T>& ta, V3<T> hb, const Pose<T>& tb, Contact<T>& c,
            Choice* ch = nullptr)
{
    V3<T> n;
    T pen;
    int sel = 0;
    if (!sat(ha, ta, hb, tb, n, pen, &sel))
        return false;
    V3<T> centerDelta = sub(tb.p, ta.p);
    const bool flip = dot(centerDelta, n) < T(0);
    if (flip)
        n = neg(n);
    if (ch) {
        ch->sel = sel;
        ch->flip = flip;
    }
    c.normal = n;
    V3<T> ref = absOf(n.x) < T(0.9f) ? V3<T>{T(1), T(0), T(0)} : V3<T>{T(0), T(1), T(0)};
    V3<T> t1 = normalize(cross(n, ref));
    V3<T> t2 = cross(n, t1);
    T nc = (dot(ta.p, n) + projectedRadius(ha, ta, n) + dot(tb.p, n) - projectedRadius(hb, tb, n)) *
           T(0.5f);
    V3<T> fallback = add(add(scale(n, nc), scale(t1, overlapMid(ha, ta, hb, tb, t1))),
                         scale(t2, overlapMid(ha, ta, hb, tb, t2)));
    V3<T> pts[4];
    T depths[4];
    uint32_t count = 0;
    auto faceCandidates = [&](V3<T> hs, const Pose<T>& ts, V3<T> ho, const Pose<T>& to,
                              V3<T> dir, int side) {
        T otherSurface = dot(boxSupport(ho, to, neg(dir)), dir);
        if (ch) {
            V3<T> l = rotate(conj(to.q), neg(dir));
            ch->support[side] = (l.x >= T(0) ? 1 : 0) | (l.y >= T(0) ? 2 : 0) | (l.z >= T(0) ? 4 : 0);
        }
        V3<T> verts[4];
        supportFace(hs, ts, dir, verts);
        for (int k = 0; k < 4; ++k) {
            if (!insideBox(verts[k], ho, to))
                continue;
            T excess = dot(verts[k], dir) - otherSurface;
            T d = maxOf(excess, T(0));
            V3<T> cand = sub(verts[k], scale(dir, d * T(0.5f)));
            if (count < 4 && !duplicate(pts, count, cand)) {
                if (ch) {
                    ch->side[count] = side;
                    ch->corner[count] = k;
                    ch->zero[count] = excess < T(0);
                }
                depths[count] = d;
                pts[count++] = cand;
            }
        }
    };
    faceCandidates(ha, ta, hb, tb, n, 0);
    faceCandidates(hb, tb, ha, ta, neg(n), 1);
    if (count == 0) {
        depths[0] = pen;
        pts[0] = fallback;
        count = 1;
        if (ch)
            ch->fallback = 1;
    }
    if (ch)
        ch->count = int(count);
    c.count = count;
    for (uint32_t k = 0; k < count; ++k) {
        c.depth[k] = depths[k];
        setAnchors(c, k, pts[k], ta, tb);
    }
    return true;
}

// ---- GJK / EPA over support maps (gjk.cpp, epa.cpp).
template <typename T>
struct Support {
    const Shape<T>* s;
    const Pose<T>* t;
    V3<T> operator()(V3<T> d) const
    {
        if (s->type == kSphere)
            return add(t->p, scale(normalize(d), s->data.x));
        if (s->type == kBox)
            return boxSupport(s->data, *t, d);
        V3<T> l = rotate(conj(t->q), d);
        uint32_t best = 0;
        T bestDot = dot(s->points[0], l);
        for (uint32_t k = 1; k < s->pointCount; ++k) {
            T v = dot(s->points[k], l);
            if (v > bestDot) {
                bestDot = v;
                best = k;
            }
        }
        return toWorld(s->points[best], *t);
    }
};

template <typename T>
struct Simplex {
    V3<T> p[4];
    uint32_t size = 0;
};

template <typename T>
inline T lsq(V3<T> v) { return dot(v, v); }

// setLineDirection, gjk.cpp:21-41.
template <typename T>
void lineCase(Simplex<T>& s, V3<T>& dir)
{
    V3<T> a = s.p[0], b = s.p[1];
    V3<T> ao = neg(a);
    V3<T> ab = sub(b, a);
    if (dot(ab, ao) > T(0)) {
        dir = cross(cross(ab, ao), ab);
        if (lsq(dir) < T(1e-12f))
            dir = cross(ab, V3<T>{T(1), T(0), T(0)});
        if (lsq(dir) < T(1e-12f))
            dir = cross(ab, V3<T>{T(0), T(1), T(0)});
    } else {
        s.size = 1;
        dir = ao;
    }
}

// setTriangleDirection, gjk.cpp:43-74.  Taking the line branch leaves the
// simplex size at 3, as the reference does.
template <typename T>
void triangleCase(Simplex<T>& s, V3<T>& dir)
{
    V3<T> a = s.p[0], b = s.p[1], c = s.p[2];
    V3<T> ao = neg(a);
    V3<T> ab = sub(b, a), ac = sub(c, a);
    V3<T> nrm = cross(ab, ac);
    if (dot(cross(nrm, ac), ao) > T(0)) {
        if (dot(ac, ao) > T(0)) {
            s.p[1] = c;
            s.size = 2;
            dir = cross(cross(ac, ao), ac);
            return;
        }
        lineCase(s, dir);
        return;
    }
    if (dot(cross(ab, nrm), ao) > T(0)) {
        lineCase(s, dir);
        return;
    }
    if (dot(nrm, ao) > T(0)) {
        dir = nrm;
    } else {
        std::swap(s.p[1], s.p[2]);
        dir = neg(nrm);
    }
}

// containsOrigin, gjk.cpp:76-124.  In the third face test the reference
// writes points[1] = D and then reads points[1] (through a reference bound
// to it) for points[2], so both become D; replayed here.
template <typename T>
bool tetraCase(Simplex<T>& s, V3<T>& dir)
{
    if (s.size == 1) {
        dir = neg(s.p[0]);
        return false;
    }
    if (s.size == 2) {
        lineCase(s, dir);
        return false;
    }
    if (s.size == 3) {
        triangleCase(s, dir);
        return false;
    }
    V3<T> a = s.p[0];
    V3<T> ao = neg(a);
    auto outside = [&](V3<T> fb, V3<T> fc, V3<T> opp) {
        V3<T> nrm = cross(sub(fb, a), sub(fc, a));
        if (dot(nrm, sub(opp, a)) > T(0))
            nrm = neg(nrm);
        return dot(nrm, ao) > T(0);
    };
    if (outside(s.p[1], s.p[2], s.p[3])) {
        s.size = 3;
        triangleCase(s, dir);
        return false;
    }
    if (outside(s.p[2], s.p[3], s.p[1])) {
        s.p[1] = s.p[2];
        s.p[2] = s.p[3];
        s.size = 3;
        triangleCase(s, dir);
        return false;
    }
    if (outside(s.p[3], s.p[1], s.p[2])) {
        s.p[1] = s.p[3];
        s.p[2] = s.p[1];
        s.size = 3;
        triangleCase(s, dir);
        return false;
    }
    return true;
}

// intersectGjk, gjk.cpp:128-158.
template <typename T>
bool gjk(const Support<T>& sa, const Support<T>& sb, Simplex<T>& s, uint32_t iterations)
{
    s = Simplex<T>{};
    V3<T> dir{T(1), T(0), T(0)};
    for (uint32_t it = 0; it < iterations; ++it) {
        ++t_gjk_iterations;
        if (lsq(dir) < T(1e-12f))
            break;
        V3<T> p = sub(sa(dir), sb(neg(dir)));
        if (dot(p, dir) <= T(0))
            break;
        for (uint32_t k = std::min(s.size, 3u); k > 0; --k)
            s.p[k] = s.p[k - 1];
        s.p[0] = p;
        s.size = std::min(s.size + 1, 4u);
        if (tetraCase(s, dir))
            return true;
    }
    return false;
}

template <typename T>
struct Face {
    uint32_t a, b, c;
    V3<T> n;
    T d;
};

// makeFace, epa.cpp:29-47.
template <typename T>
Face<T> makeFace(const std::vector<V3<T>>& v, uint32_t a, uint32_t b, uint32_t c)
{
    V3<T> n = cross(sub(v[b], v[a]), sub(v[c], v[a]));
    T len = length(n);
    if (len > T(1e-8f))
        n = scale(n, T(1) / len);
    T d = dot(n, v[a]);
    if (d < T(0)) {
        std::swap(b, c);
        n = neg(n);
        d = -d;
    }
    return {a, b, c, n, d};
}

// Peak polytope sizes seen by epa() (faces, horizon edges, vertices), for
// sizing bounded GPU storage.
std::atomic<uint32_t> g_peak_faces{0}, g_peak_edges{0}, g_peak_vertices{0};
// Histogram of EPA iterations per call (index = iterations run, 0..64) and of
// visible-face counts per iteration (0..255, capped).
std::atomic<uint64_t> g_epa_iters[66], g_epa_visible[257];
inline void note_peak(std::atomic<uint32_t>& a, size_t v)
{
    uint32_t cur = a.load();
    while (v > cur && !a.compare_exchange_weak(cur, uint32_t(v))) {
    }
}

// computeEpa (epa.cpp:64-130) over a polytope of distinct faces with
// multiplicities.  The reference's face list can hold many copies of one face
// (a horizon edge that survives |s| times gives |s| copies of its new face);
// all copies share their visibility, so the list's evolution is fixed by:
//   - the closest face: the first minimum in list order;
//   - the horizon: for each unordered vertex pair the calls of
//     addHorizonEdge (epa.cpp:49-60) in its two directions form a running sum
//     s of +1 / -1, and the list ends holding |s| copies of the direction
//     sign(s) (a call against the held direction erases the oldest copy); s
//     is a sum, so it does not depend on the order of the calls;
//   - the new faces: one per surviving directed edge, |s| copies each.
// Only the closest-face choice needs the order of the copies.  Distinct faces
// are kept in the order of their first copies: faces of older iterations come
// first (kept faces keep their order, new ones are appended) and the initial
// four are in the reference's order.  Within a later iteration, the order of
// the new faces is the order of the first surviving copy of each directed
// edge in the horizon list; it is known exactly when every visible face had
// one copy and a known order (then the horizon list is replayed as the
// reference builds it) and unknown otherwise.  A minimum shared by faces of
// one iteration whose order is unknown, with different normals, is reported
// as undecided (-2); every other run equals the reference's.
struct MultiStats {
    uint32_t peakFaces = 0, peakEdges = 0;
    uint64_t peakMultiplicity = 0;
    uint32_t iterations = 0;
};

template <typename T>
struct MFace {
    Face<T> f;
    uint64_t mult;
    uint32_t gen;  // iteration that made it (0: the faces the run started from)
};

// Runs iterations [first, opt.epaIterations) over the distinct faces `faces`
// (first-copy order, every face of generation 0 in known order) and vertices
// `v`.  Returns 1 converged, 0 the reference's failure, -1 more than
// opt.multiFaces distinct faces, -2 undecided closest face.  With
// opt.multiExactOrder the order of an iteration's new faces is derived where
// it is known (see above); without it every later iteration's order counts as
// unknown.
template <typename T>
int epaMultiRun(const Support<T>& sa, const Support<T>& sb, const NarrowOptions& opt,
                std::vector<V3<T>>& v, std::vector<MFace<T>>& faces, uint32_t first,
                V3<T>& normal, T& depth, MultiStats* ms)
{
    const uint32_t iterations = opt.epaIterations;
    const T tol = T(opt.epaTolerance);
    std::vector<uint8_t> ordered(iterations + 2, 0);  // order of each generation's faces known
    ordered[0] = 1;
    const uint32_t V = 4 + iterations;
    std::vector<int64_t> sum(size_t(V) * V);
    std::vector<uint32_t> touched;
    struct Edge {
        uint32_t a, b;
    };
    for (uint32_t it = first; it < iterations; ++it) {
        if (ms)
            ms->iterations = it + 1;
        if (faces.empty())
            return 0;
        size_t best = 0;
        for (size_t f = 1; f < faces.size(); ++f)
            if (faces[f].f.d < faces[best].f.d)
                best = f;
        const uint32_t g0 = faces[best].gen;
        if (!ordered[g0]) {
            for (size_t f = best + 1; f < faces.size() && faces[f].gen == g0; ++f) {
                const Face<T>& o = faces[f].f;
                if (o.d == faces[best].f.d &&
                    (o.n.x != faces[best].f.n.x || o.n.y != faces[best].f.n.y ||
                     o.n.z != faces[best].f.n.z))
                    return -2;
            }
        }
        Face<T> cf = faces[best].f;
        V3<T> sup = sub(sa(cf.n), sb(neg(cf.n)));
        T sd = dot(sup, cf.n);
        if (sd - cf.d <= tol) {
            normal = cf.n;
            depth = maxOf(cf.d, T(0));
            return 1;
        }
        uint32_t si = uint32_t(v.size());
        v.push_back(sup);
        std::vector<MFace<T>> keep, added;
        std::vector<const MFace<T>*> visible;
        bool exact = opt.multiExactOrder;
        for (const MFace<T>& mf : faces) {
            if (dot(mf.f.n, sub(sup, v[mf.f.a])) > tol) {
                visible.push_back(&mf);
                exact = exact && mf.mult == 1 && ordered[mf.gen];
            } else {
                keep.push_back(mf);
            }
        }
        const uint32_t gen = it + 1 - first + 1;  // generations after the start are 2, 3, ...
        if (exact) {
            // The reference's horizon list itself (one copy of each face).
            std::vector<Edge> horizon;
            auto addEdge = [&](uint32_t a, uint32_t b) {
                for (size_t e = 0; e < horizon.size(); ++e)
                    if (horizon[e].a == b && horizon[e].b == a) {
                        horizon.erase(horizon.begin() + e);
                        return;
                    }
                horizon.push_back({a, b});
            };
            for (const MFace<T>* mf : visible) {
                addEdge(mf->f.a, mf->f.b);
                addEdge(mf->f.b, mf->f.c);
                addEdge(mf->f.c, mf->f.a);
            }
            for (const Edge& e : horizon) {
                bool merged = false;
                for (MFace<T>& m : added)  // a repeated directed edge: one more copy
                    if (m.f.a == e.a && m.f.b == e.b) {
                        ++m.mult;
                        merged = true;
                        break;
                    }
                if (!merged)
                    added.push_back({makeFace(v, e.a, e.b, si), 1, gen});
            }
        } else {
            // Running sums per unordered vertex pair, weighted by multiplicity;
            // new faces in the order of each pair's first call.
            for (const MFace<T>* mf : visible) {
                const Face<T>& f = mf->f;
                const uint32_t e[3][2] = {{f.a, f.b}, {f.b, f.c}, {f.c, f.a}};
                for (auto& ed : e) {
                    uint32_t lo = std::min(ed[0], ed[1]), hi = std::max(ed[0], ed[1]);
                    size_t key = size_t(lo) * V + hi;
                    bool fresh = true;
                    for (uint32_t t : touched)
                        fresh = fresh && t != key;
                    if (fresh)
                        touched.push_back(uint32_t(key));
                    const int64_t m = int64_t(mf->mult);
                    sum[key] += ed[0] < ed[1] ? m : -m;
                }
            }
            for (uint32_t key : touched) {
                int64_t sk = sum[key];
                sum[key] = 0;
                if (sk == 0)
                    continue;
                uint32_t lo = key / V, hi = key % V;
                uint32_t a = sk > 0 ? lo : hi, b = sk > 0 ? hi : lo;
                added.push_back({makeFace(v, a, b, si), uint64_t(sk > 0 ? sk : -sk), gen});
            }
            touched.clear();
        }
        ordered[gen] = exact ? 1 : 0;
        if (added.empty())
            return 0;
        if (opt.multiFaces && keep.size() + added.size() > opt.multiFaces)
            return -1;
        faces = std::move(keep);
        faces.insert(faces.end(), added.begin(), added.end());
        if (ms) {
            ms->peakEdges = std::max<uint32_t>(ms->peakEdges, uint32_t(added.size()));
            ms->peakFaces = std::max<uint32_t>(ms->peakFaces, uint32_t(faces.size()));
            for (const MFace<T>& f : faces)
                ms->peakMultiplicity = std::max(ms->peakMultiplicity, f.mult);
        }
    }
    return 0;
}

// The multiplicity form from the start (four initial faces).
template <typename T>
int epaMulti(const Support<T>& sa, const Support<T>& sb, const Simplex<T>& s,
             const NarrowOptions& opt, V3<T>& normal, T& depth, MultiStats* ms)
{
    normal = {T(0), T(0), T(0)};
    depth = T(0);
    if (s.size < 4 || opt.epaIterations == 0)
        return 0;
    std::vector<V3<T>> v(s.p, s.p + 4);
    std::vector<MFace<T>> faces = {{makeFace(v, 0, 1, 2), 1, 0}, {makeFace(v, 0, 3, 1), 1, 0},
                                   {makeFace(v, 0, 2, 3), 1, 0}, {makeFace(v, 1, 3, 2), 1, 0}};
    return epaMultiRun(sa, sb, opt, v, faces, 0, normal, depth, ms);
}

// Continuation of the bounded exact EPA (epa below) at iteration `it`, where
// its polytope outgrew opt.maxFaces / opt.maxEdges: the exact list `list`
// (before that iteration) becomes distinct faces in first-copy order with
// their copy counts, all in known order, and the run goes on from `it`.
template <typename T>
int epaContinue(const Support<T>& sa, const Support<T>& sb, const NarrowOptions& opt,
                std::vector<V3<T>>& v, const std::vector<Face<T>>& list, uint32_t it,
                V3<T>& normal, T& depth, MultiStats* ms)
{
    std::vector<MFace<T>> faces;
    for (const Face<T>& f : list) {
        bool merged = false;
        for (MFace<T>& m : faces)
            if (m.f.a == f.a && m.f.b == f.b && m.f.c == f.c) {
                ++m.mult;
                merged = true;
                break;
            }
        if (!merged)
            faces.push_back({f, 1, 0});
    }
    if (opt.multiFaces && faces.size() > opt.multiFaces)
        return -1;
    return epaMultiRun(sa, sb, opt, v, faces, it, normal, depth, ms);
}

// computeEpa, epa.cpp:64-130.  Returns 1 (converged), 0 (the reference's
// failure), -1 (a bound of opt was exceeded; never with the defaults) or,
// with opt.multiFaces, what the multiplicity continuation returns from the
// iteration where a bound was exceeded (1, 0, -1 its own bound, -2 undecided).
template <typename T>
int epa(const Support<T>& sa, const Support<T>& sb, const Simplex<T>& s, const NarrowOptions& opt,
        V3<T>& normal, T& depth)
{
    const uint32_t iterations = opt.epaIterations;
    const T tol = T(opt.epaTolerance);
    normal = {T(0), T(0), T(0)};
    depth = T(0);
    if (s.size < 4 || iterations == 0)
        return 0;
    std::vector<V3<T>> v(s.p, s.p + 4);
    std::vector<Face<T>> faces = {makeFace(v, 0, 1, 2), makeFace(v, 0, 3, 1),
                                  makeFace(v, 0, 2, 3), makeFace(v, 1, 3, 2)};
    struct Edge {
        uint32_t a, b;
    };
    struct IterNote {
        uint32_t it = 0;
        ~IterNote() { g_epa_iters[it < 65 ? it : 65]++; }
    } note;
    for (uint32_t it = 0; it < iterations; ++it, ++note.it) {
        ++t_epa_iterations;
        if (faces.empty())
            return 0;
        size_t best = 0;
        for (size_t f = 1; f < faces.size(); ++f)
            if (faces[f].d < faces[best].d)
                best = f;
        Face<T> cf = faces[best];
        V3<T> sup = sub(sa(cf.n), sb(neg(cf.n)));
        T sd = dot(sup, cf.n);
        if (sd - cf.d <= tol) {
            normal = cf.n;
            depth = maxOf(cf.d, T(0));
            return 1;
        }
        uint32_t si = uint32_t(v.size());
        v.push_back(sup);
        std::vector<Edge> horizon;
        std::vector<Face<T>> keep;
        bool overflow = false;
        auto addEdge = [&](uint32_t a, uint32_t b) {
            for (size_t e = 0; e < horizon.size(); ++e)
                if (horizon[e].a == b && horizon[e].b == a) {
                    horizon.erase(horizon.begin() + e);
                    return;
                }
            if (opt.maxEdges && horizon.size() >= opt.maxEdges)
                overflow = true;
            horizon.push_back({a, b});
        };
        for (const Face<T>& f : faces) {
            if (dot(f.n, sub(sup, v[f.a])) > tol) {
                addEdge(f.a, f.b);
                addEdge(f.b, f.c);
                addEdge(f.c, f.a);
            } else {
                keep.push_back(f);
            }
        }
        note_peak(g_peak_edges, horizon.size());
        note_peak(g_peak_vertices, v.size());
        {
            size_t vis = faces.size() - keep.size();
            g_epa_visible[vis < 256 ? vis : 256]++;
        }
        if (!overflow && horizon.empty())
            return 0;
        if (overflow || (opt.maxFaces && keep.size() + horizon.size() > opt.maxFaces)) {
            if (!opt.multiFaces)
                return -1;
            v.pop_back();  // the continuation takes this iteration again
            return epaContinue(sa, sb, opt, v, faces, it, normal, depth,
                               static_cast<MultiStats*>(nullptr));
        }
        faces = std::move(keep);
        for (const Edge& e : horizon)
            faces.push_back(makeFace(v, e.a, e.b, si));
        note_peak(g_peak_faces, faces.size());
    }
    return 0;
}

// Planted mistakes for the tests of the tests (orc_set_test_mutation): 0 none,
// 1 the hull contact point from the supports along the wrong signs, 2 the
// contact point at A's support point (weight 1 instead of 1/2 on each side).
std::atomic<int> g_mutation{0};

// Hull-involving pairs: GJK then EPA; one contact point at the midpoint of
// the two deepest support points along the EPA normal (A toward B).
template <typename T>
bool convexPair(const Shape<T>& a, const Pose<T>& ta, const Shape<T>& b, const Pose<T>& tb,
                const NarrowOptions& opt, Contact<T>& c, int* status)
{
    Support<T> sa{&a, &ta}, sb{&b, &tb};
    Simplex<T> s;
    const uint64_t flops0 = t_flops.total();
    const bool intersect = gjk(sa, sb, s, opt.gjkIterations);
    t_gjk_flops += t_flops.total() - flops0;
    if (!intersect) {
        if (status)
            *status = 0;
        return false;
    }
    V3<T> n;
    T depth;
    const int st = epa(sa, sb, s, opt, n, depth);
    if (st != 1) {
        if (status)
            *status = st == 0 ? 2 : 3;
        return false;
    }
    if (status)
        *status = 1;
    c.normal = n;
    c.count = 1;
    c.depth[0] = depth;
    V3<T> w = scale(add(sa(n), sb(neg(n))), T(0.5f));
    const int mutation = g_mutation.load(std::memory_order_relaxed);
    if (mutation == 1)
        w = scale(add(sa(neg(n)), sb(n)), T(0.5f));
    else if (mutation == 2)
        w = sa(n);
    setAnchors(c, 0, w, ta, tb);
    return true;
}

} // namespace np

template <typename T>
bool generateContact(const Shape<T>& a, const Pose<T>& ta, const Shape<T>& b, const Pose<T>& tb,
                     const NarrowOptions& opt, Contact<T>& c, int* gjkStatus)
{
    c = Contact<T>{};
    if (a.type == kHull || b.type == kHull)
        return np::convexPair(a, ta, b, tb, opt, c, gjkStatus);
    if (a.type == kSphere && b.type == kSphere)
        return np::sphereSphere(a.data.x, ta, b.data.x, tb, c);
    if (a.type == kSphere && b.type == kBox)
        return np::sphereBox(a.data.x, ta, b.data, tb, c);
    if (a.type == kBox && b.type == kSphere) {
        if (!np::sphereBox(b.data.x, tb, a.data, ta, c))
            return false;
        c.normal = neg(c.normal);
        for (uint32_t k = 0; k < c.count; ++k)
            std::swap(c.anchorA[k], c.anchorB[k]);
        return true;
    }
    return np::boxBox(a.data, ta, b.data, tb, c);
}

// FP32 with every arithmetic operation counted (add / subtract, multiply,
// divide, square root; negation, comparison and absolute value are not
// floating-point work): each operation is the float operation, so values and
// branches are those of generateContact<float>.
struct Counted {
    float v;
    Counted() : v(0.0f) {}
    template <typename U, typename = std::enable_if_t<std::is_arithmetic_v<U>>>
    Counted(U x) : v(float(x))
    {
    }
    explicit operator float() const { return v; }
};
inline Counted operator+(Counted a, Counted b)
{
    ++t_flops.add;
    return Counted(a.v + b.v);
}
inline Counted operator-(Counted a, Counted b)
{
    ++t_flops.add;
    return Counted(a.v - b.v);
}
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted operator*(Counted a, Counted b)
{
    ++t_flops.mul;
    return Counted(a.v * b.v);
}
inline Counted operator/(Counted a, Counted b)
{
    ++t_flops.div;
    return Counted(a.v / b.v);
}
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }
template <>
inline Counted absOf<Counted>(Counted x) { return Counted(std::abs(x.v)); }
template <>
inline Counted sqrtOf<Counted>(Counted x)
{
    ++t_flops.sqrt;
    return Counted(std::sqrt(x.v));
}
template <>
inline Counted bigOf<Counted>() { return Counted(std::numeric_limits<float>::max()); }

template bool generateContact<Counted>(const Shape<Counted>&, const Pose<Counted>&,
                                       const Shape<Counted>&, const Pose<Counted>&,
                                       const NarrowOptions&, Contact<Counted>&, int*);
template bool generateContact<float>(const Shape<float>&, const Pose<float>&, const Shape<float>&,
                                     const Pose<float>&, const NarrowOptions&, Contact<float>&,
                                     int*);
template bool generateContact<double>(const Shape<double>&, const Pose<double>&,
                                      const Shape<double>&, const Pose<double>&,
                                      const NarrowOptions&, Contact<double>&, int*);

} // namespace orc

namespace orc {

// ---------------------------------------------------------------- FP64 adjoint
// Forward-mode dual numbers over the same templated contact code: 14 tangent
// directions (position and quaternion of A, then of B), so one evaluation
// gives the Jacobian of the normal, depths and world points with respect to
// both poses.  The orientation enters as q/|q|, so quaternion derivatives are
// those of the contact as a function of the normalised orientation.
struct Dual {
    static constexpr int N = 14;
    double v;
    double d[N];
    Dual() : v(0.0) { std::fill(d, d + N, 0.0); }
    Dual(double x) : v(x) { std::fill(d, d + N, 0.0); }
};
inline Dual operator+(const Dual& a, const Dual& b)
{
    Dual r(a.v + b.v);
    for (int i = 0; i < Dual::N; ++i)
        r.d[i] = a.d[i] + b.d[i];
    return r;
}
inline Dual operator-(const Dual& a, const Dual& b)
{
    Dual r(a.v - b.v);
    for (int i = 0; i < Dual::N; ++i)
        r.d[i] = a.d[i] - b.d[i];
    return r;
}
inline Dual operator-(const Dual& a)
{
    Dual r(-a.v);
    for (int i = 0; i < Dual::N; ++i)
        r.d[i] = -a.d[i];
    return r;
}
inline Dual operator*(const Dual& a, const Dual& b)
{
    Dual r(a.v * b.v);
    for (int i = 0; i < Dual::N; ++i)
        r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
inline Dual operator/(const Dual& a, const Dual& b)
{
    Dual r(a.v / b.v);
    for (int i = 0; i < Dual::N; ++i)
        r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
template <>
inline Dual absOf<Dual>(Dual x) { return x.v < 0.0 ? -x : x; }
template <>
inline Dual sqrtOf<Dual>(Dual x)
{
    Dual r(std::sqrt(x.v));
    for (int i = 0; i < Dual::N; ++i)
        r.d[i] = r.v > 0.0 ? x.d[i] / (2.0 * r.v) : 0.0;
    return r;
}
template <>
inline Dual bigOf<Dual>() { return Dual(std::numeric_limits<double>::max()); }

// Sphere-sphere, sphere-box, box-sphere, box-box (NarrowPhase::generateContact,
// narrowphase.cpp:326-374); hull pairs are not differentiated.
template <typename T>
bool contactNoHull(int typeA, V3<T> dataA, const Pose<T>& ta, int typeB, V3<T> dataB,
                   const Pose<T>& tb, Contact<T>& c, Choice* ch)
{
    c.count = 0;
    if (typeA == kSphere && typeB == kSphere)
        return np::sphereSphere(dataA.x, ta, dataB.x, tb, c);
    if (typeA == kSphere && typeB == kBox)
        return np::sphereBox(dataA.x, ta, dataB, tb, c, ch);
    if (typeA == kBox && typeB == kSphere) {
        if (!np::sphereBox(dataB.x, tb, dataA, ta, c, ch))
            return false;
        c.normal = neg(c.normal);
        return true;
    }
    if (typeA == kBox && typeB == kBox)
        return np::boxBox(dataA, ta, dataB, tb, c, ch);
    return false;
}

template <typename T>
Pose<T> normalisedPose(const double p[7])
{
    Pose<T> t{{T(p[0]), T(p[1]), T(p[2])}, {T(p[3]), T(p[4]), T(p[5]), T(p[6])}};
    T len = sqrtOf(t.q.x * t.q.x + t.q.y * t.q.y + t.q.z * t.q.z + t.q.w * t.q.w);
    t.q = {t.q.x / len, t.q.y / len, t.q.z / len, t.q.w / len};
    return t;
}

// Pose with tangent directions base..base+6 on its 7 components.
inline Pose<Dual> normalisedPoseSeeded(const double p[7], int base)
{
    Dual c[7];
    for (int q = 0; q < 7; ++q) {
        c[q] = Dual(p[q]);
        c[q].d[base + q] = 1.0;
    }
    Pose<Dual> t{{c[0], c[1], c[2]}, {c[3], c[4], c[5], c[6]}};
    Dual len = sqrtOf(t.q.x * t.q.x + t.q.y * t.q.y + t.q.z * t.q.z + t.q.w * t.q.w);
    t.q = {t.q.x / len, t.q.y / len, t.q.z / len, t.q.w / len};
    return t;
}

// Upstream gradients of one manifold (SoA planes of m entries at index k).
struct Upstream {
    double n[3], d[4], p[4][3];
    Upstream(const float* gn, const float* gd, const float* gp, uint64_t m, uint64_t k)
    {
        for (int a = 0; a < 3; ++a)
            n[a] = gn ? gn[a * m + k] : 0.0;
        for (int j = 0; j < 4; ++j) {
            d[j] = gd ? gd[j * m + k] : 0.0;
            for (int a = 0; a < 3; ++a)
                p[j][a] = gp ? gp[(3 * j + a) * m + k] : 0.0;
        }
    }
    // Linearised loss: g . (normal, depths, points) of a contact.
    template <typename T>
    T phi(const Contact<T>& c) const
    {
        T s = T(n[0]) * c.normal.x + T(n[1]) * c.normal.y + T(n[2]) * c.normal.z;
        for (uint32_t j = 0; j < c.count && j < 4; ++j)
            s = s + T(d[j]) * c.depth[j] + T(p[j][0]) * c.world[j].x +
                T(p[j][1]) * c.world[j].y + T(p[j][2]) * c.world[j].z;
        return s;
    }
};

} // namespace orc

// ====================================================================== C ABI
using namespace orc;

namespace {

thread_local std::string lastError;

template <typename T>
struct View {
    const physgpu_colliders* c;
    std::vector<V3<T>> hull;
    explicit View(const physgpu_colliders* cc) : c(cc)
    {
        const uint32_t H = c->hull_points ? c->hull_point_count : 0;
        hull.resize(H);
        for (uint32_t k = 0; k < H; ++k)
            hull[k] = {T(c->hull_points[k]), T(c->hull_points[H + k]),
                       T(c->hull_points[2 * H + k])};
    }
    Pose<T> pose(uint32_t i) const
    {
        const uint32_t n = c->count;
        const float* p = c->pose;
        return {{T(p[i]), T(p[n + i]), T(p[2 * n + i])},
                {T(p[3 * n + i]), T(p[4 * n + i]), T(p[5 * n + i]), T(p[6 * n + i])}};
    }
    Shape<T> shape(uint32_t i) const
    {
        const uint32_t n = c->count;
        Shape<T> s{};
        s.type = c->shape_type[i];
        s.data = {T(c->shape_data[i]), T(c->shape_data[n + i]), T(c->shape_data[2 * n + i])};
        if (s.type == kHull) {
            if (!c->hull_range)
                throw std::invalid_argument("hull collider without hull_range");
            uint32_t first = c->hull_range[i], count = c->hull_range[n + i];
            if (count == 0 || size_t(first) + count > hull.size())
                throw std::invalid_argument("hull_range outside hull_points");
            s.points = hull.data() + first;
            s.pointCount = count;
        } else if (s.type != kSphere && s.type != kBox) {
            throw std::invalid_argument("unknown shape type");
        }
        return s;
    }
};

uint32_t g_max_faces = 0, g_max_edges = 0, g_multi_faces = 0;
bool g_multi_exact = true;

NarrowOptions makeOptions(uint32_t gjkIt, uint32_t epaIt, float tol)
{
    NarrowOptions o;
    o.maxFaces = g_max_faces;
    o.maxEdges = g_max_edges;
    o.multiFaces = g_multi_faces;
    o.multiExactOrder = g_multi_exact;
    if (gjkIt)
        o.gjkIterations = gjkIt;
    if (epaIt)
        o.epaIterations = epaIt;
    if (tol > 0.0f)
        o.epaTolerance = tol;
    return o;
}

uint32_t pairClass(int ta, int tb)
{
    if (ta > tb)
        std::swap(ta, tb);
    if (tb < 2)
        return uint32_t(ta + tb);             // SS 0, SB 1, BB 2
    return uint32_t(3 + ta);                  // SH 3, BH 4, HH 5
}

void store(const Contact<float>& c, physgpu_manifold* m)
{
    m->normal[0] = c.normal.x;
    m->normal[1] = c.normal.y;
    m->normal[2] = c.normal.z;
    m->point_count = c.count;
    for (int k = 0; k < 4; ++k) {
        m->anchor_a[k][0] = c.anchorA[k].x;
        m->anchor_a[k][1] = c.anchorA[k].y;
        m->anchor_a[k][2] = c.anchorA[k].z;
        m->anchor_b[k][0] = c.anchorB[k].x;
        m->anchor_b[k][1] = c.anchorB[k].y;
        m->anchor_b[k][2] = c.anchorB[k].z;
        m->penetration[k] = c.depth[k];
    }
}

int fail(const std::exception& e)
{
    lastError = e.what();
    return PHYSGPU_INVALID_ARGUMENT;
}

void findPairs(const physgpu_aabb* bounds, uint32_t n, int algorithm, uint32_t workers,
               std::vector<Pair>& pairs, physgpu_broadphase_stats& st)
{
    std::memset(&st, 0, sizeof(st));
    if (algorithm == PHYSGPU_UNIFORM_GRID) {
        uniformGrid(bounds, n, pairs, st);
        return;
    }
    uint64_t tests = 0;
    sweepAndPrune(bounds, n, workers, pairs, tests);
    st.aabb_pairs = pairs.size();
    if (algorithm == PHYSGPU_SWEEP_AND_PRUNE)
        st.x_window_comparisons = tests;
    else
        st.tree_proxy_count = n;
}

} // namespace

// Contact of pair (a, b) in scalar type T (double or Dual): sphere, box and
// the polyhedral hull pairs (box-hull, hull-hull through GJK + EPA); false
// for sphere-hull pairs, which the adjoint does not differentiate.
template <typename T>
static bool contactTyped(const View<double>& v, uint32_t a, uint32_t b, const Pose<T>& ta,
                         const Pose<T>& tb, Contact<T>& c, Choice* ch)
{
    const Shape<double> sa = v.shape(a), sb = v.shape(b);
    std::vector<V3<T>> pa, pb;
    auto convert = [](const Shape<double>& s, std::vector<V3<T>>& buf) {
        Shape<T> o{};
        o.type = s.type;
        o.data = {T(s.data.x), T(s.data.y), T(s.data.z)};
        if (s.type == kHull) {
            buf.resize(s.pointCount);
            for (uint32_t k = 0; k < s.pointCount; ++k)
                buf[k] = {T(s.points[k].x), T(s.points[k].y), T(s.points[k].z)};
            o.points = buf.data();
            o.pointCount = s.pointCount;
        }
        return o;
    };
    const Shape<T> A = convert(sa, pa), B = convert(sb, pb);
    c.count = 0;
    if (A.type == kHull || B.type == kHull) {
        if (A.type == kSphere || B.type == kSphere)
            return false;
        return np::convexPair(A, ta, B, tb, makeOptions(64, 64, 1e-6f), c, nullptr);
    }
    return contactNoHull(A.type, A.data, ta, B.type, B.data, tb, c, ch);
}

static bool sphereHull(const View<double>& v, uint32_t a, uint32_t b)
{
    const int ta = v.shape(a).type, tb = v.shape(b).type;
    return (ta == kHull && tb == kSphere) || (ta == kSphere && tb == kHull);
}

// FP64 contact of pair k at poses pa, pb (7 doubles each, q normalised).
static bool contactF64(View<double>& v, uint32_t a, uint32_t b, const double pa[7],
                       const double pb[7], Contact<double>& c, Choice* ch)
{
    return contactTyped(v, a, b, normalisedPose<double>(pa), normalisedPose<double>(pb), c, ch);
}

static void posesOf(const physgpu_colliders* c, uint32_t i, double out[7])
{
    const uint32_t n = c->count;
    for (int q = 0; q < 7; ++q)
        out[q] = c->pose[q * n + i];
}

extern "C" {

const char* orc_last_error() { return lastError.c_str(); }

// EPA iteration histogram (66 counters) and visible-face histogram (257).
void orc_epa_histograms(uint64_t* iters, uint64_t* visible, int reset)
{
    for (int i = 0; i < 66; ++i) {
        iters[i] = np::g_epa_iters[i].load();
        if (reset)
            np::g_epa_iters[i] = 0;
    }
    for (int i = 0; i < 257; ++i) {
        visible[i] = np::g_epa_visible[i].load();
        if (reset)
            np::g_epa_visible[i] = 0;
    }
}

// Test hook: plant one of the mistakes of g_mutation (0 restores the oracle).
void orc_set_test_mutation(int mutation) { np::g_mutation.store(mutation); }

// EPA polytope bounds used by the following calls (0, 0 = unbounded, the
// reference's behaviour).  Test infrastructure: not thread-safe.
void orc_set_epa_bounds(uint32_t maxFaces, uint32_t maxEdges)
{
    g_max_faces = maxFaces;
    g_max_edges = maxEdges;
}

// Continuation past the bounds in multiplicity form with at most
// `multiFaces` distinct faces (0: none, the bounded run stops with the
// overflow status); exactOrder as NarrowOptions::multiExactOrder.
void orc_set_epa_continuation(uint32_t multiFaces, int exactOrder)
{
    g_multi_faces = multiFaces;
    g_multi_exact = exactOrder != 0;
}

// Peak EPA polytope sizes since the last reset (faces, edges, vertices).
void orc_epa_peak(uint32_t* out, int reset)
{
    out[0] = np::g_peak_faces.load();
    out[1] = np::g_peak_edges.load();
    out[2] = np::g_peak_vertices.load();
    if (reset) {
        np::g_peak_faces = 0;
        np::g_peak_edges = 0;
        np::g_peak_vertices = 0;
    }
}

int orc_compute_aabbs(const physgpu_colliders* c, physgpu_aabb* out)
{
    try {
        View<float> v(c);
        for (uint32_t i = 0; i < c->count; ++i) {
            Box6<float> b = aabbOf(v.shape(i), v.pose(i));
            out[i] = {{b.mn.x, b.mn.y, b.mn.z}, {b.mx.x, b.mx.y, b.mx.z}};
        }
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// Same contract as the reference shim's ref_find_candidate_pairs.  The tree
// algorithm returns the tree's pair set (all overlapping tight bounds, sorted).
int orc_find_candidate_pairs(const physgpu_aabb* bounds, uint32_t n, int algorithm,
                             uint32_t workers, physgpu_pair* out, uint64_t capacity,
                             uint64_t* count, physgpu_broadphase_stats* stats)
{
    try {
        if (algorithm < 0 || algorithm > 2)
            throw std::invalid_argument("unknown broadphase algorithm");
        std::vector<Pair> pairs;
        physgpu_broadphase_stats st;
        findPairs(bounds, n, algorithm, workers, pairs, st);
        *count = pairs.size();
        if (out)
            for (size_t k = 0; k < pairs.size() && k < capacity; ++k)
                out[k] = {pairs[k].first, pairs[k].second};
        if (stats)
            *stats = st;
        return pairs.size() > capacity ? PHYSGPU_CAPACITY_EXCEEDED : PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// Per-pair slots (not compacted); hit[k] = contact.  gjkStatus (may be NULL)
// gets 0/1/2 for hull pairs (no hit / EPA ok / EPA failed), -1 otherwise.
int orc_generate_contacts(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                          uint32_t gjkIterations, uint32_t epaIterations, float tolerance,
                          uint32_t workers, physgpu_manifold* out, uint8_t* hit,
                          int32_t* gjkStatus)
{
    try {
        View<float> v(c);
        NarrowOptions opt = makeOptions(gjkIterations, epaIterations, tolerance);
        auto one = [&](size_t k) {
            uint32_t a = pairs[k].first, b = pairs[k].second;
            Contact<float> ct;
            int status = -1;
            Shape<float> sa = v.shape(a), sb = v.shape(b);
            bool h = generateContact(sa, v.pose(a), sb, v.pose(b), opt, ct, &status);
            physgpu_manifold& m = out[k];
            std::memset(&m, 0, sizeof(m));
            m.first = a;
            m.second = b;
            m.pair_index = uint32_t(k);
            m.shape_pair = pairClass(sa.type, sb.type);
            store(ct, &m);
            hit[k] = h ? 1 : 0;
            if (gjkStatus)
                gjkStatus[k] = status;
        };
        Pool::run(n >= 4096 ? workers : 1, n, one);
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// Floating-point operations of generateContact in FP32 (the operation
// counts of the Counted scalar), per pair k: out[k*8 + 0..7] = add/subtract,
// multiply, divide, square root, of which in GJK, then GJK iterations, EPA
// iterations, and whether the pair touches (0 / 1).
int orc_count_flops(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                    uint32_t gjkIterations, uint32_t epaIterations, float tolerance,
                    uint32_t workers, uint64_t* out)
{
    try {
        View<Counted> v(c);
        NarrowOptions opt = makeOptions(gjkIterations, epaIterations, tolerance);
        auto one = [&](size_t k) {
            const uint32_t a = pairs[k].first, b = pairs[k].second;
            t_flops = FlopCount{};
            t_gjk_iterations = t_epa_iterations = t_gjk_flops = 0;
            Contact<Counted> ct;
            const bool h = generateContact(v.shape(a), v.pose(a), v.shape(b), v.pose(b), opt, ct,
                                           nullptr);
            uint64_t* o = out + k * 8;
            o[0] = t_flops.add;
            o[1] = t_flops.mul;
            o[2] = t_flops.div;
            o[3] = t_flops.sqrt;
            o[4] = t_gjk_flops;
            o[5] = t_gjk_iterations;
            o[6] = t_epa_iterations;
            o[7] = h ? 1 : 0;
        };
        Pool::run(n >= 4096 ? workers : 1, n, one);
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// FP64 re-evaluation of the same pairs: per pair 1 + 3 + 4*7 doubles
// (point count, normal, 4 x {anchorA, anchorB, depth}).
int orc_generate_contacts_f64(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                              uint32_t gjkIterations, uint32_t epaIterations, float tolerance,
                              double* out, uint8_t* hit)
{
    try {
        View<double> v(c);
        NarrowOptions opt = makeOptions(gjkIterations, epaIterations, tolerance);
        for (uint64_t k = 0; k < n; ++k) {
            uint32_t a = pairs[k].first, b = pairs[k].second;
            Contact<double> ct;
            bool h = generateContact(v.shape(a), v.pose(a), v.shape(b), v.pose(b), opt, ct);
            hit[k] = h ? 1 : 0;
            double* o = out + k * 32;
            o[0] = ct.count;
            o[1] = ct.normal.x;
            o[2] = ct.normal.y;
            o[3] = ct.normal.z;
            for (int p = 0; p < 4; ++p) {
                o[4 + p * 7 + 0] = ct.anchorA[p].x;
                o[4 + p * 7 + 1] = ct.anchorA[p].y;
                o[4 + p * 7 + 2] = ct.anchorA[p].z;
                o[4 + p * 7 + 3] = ct.anchorB[p].x;
                o[4 + p * 7 + 4] = ct.anchorB[p].y;
                o[4 + p * 7 + 5] = ct.anchorB[p].z;
                o[4 + p * 7 + 6] = ct.depth[p];
            }
        }
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// GJK then the multiplicity form of EPA (np::epaMulti), unbounded.
// status: 0 GJK miss, 1 converged, 2 failure, 4 undecided closest face;
// stats[k*4 + 0..3] = peak distinct faces, peak distinct new faces in one
// iteration, iterations, peak multiplicity (saturated to 2^32-1).
int orc_gjk_epa_multi(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                      uint32_t gjkIterations, uint32_t epaIterations, float tolerance,
                      int32_t* status, float* normal, float* depth, uint32_t* stats)
{
    try {
        View<float> v(c);
        for (uint64_t k = 0; k < n; ++k) {
            Shape<float> sa = v.shape(pairs[k].first), sb = v.shape(pairs[k].second);
            Pose<float> ta = v.pose(pairs[k].first), tb = v.pose(pairs[k].second);
            np::Support<float> fa{&sa, &ta}, fb{&sb, &tb};
            np::Simplex<float> s;
            normal[3 * k] = normal[3 * k + 1] = normal[3 * k + 2] = 0.0f;
            depth[k] = 0.0f;
            np::MultiStats ms;
            if (!np::gjk(fa, fb, s, gjkIterations)) {
                status[k] = 0;
            } else {
                NarrowOptions opt = makeOptions(gjkIterations, epaIterations, tolerance);
                opt.maxFaces = opt.maxEdges = 0;
                V3<float> nn;
                float d;
                const int st = np::epaMulti(fa, fb, s, opt, nn, d, &ms);
                status[k] = st == 1 ? 1 : (st == 0 ? 2 : 4);
                normal[3 * k] = nn.x;
                normal[3 * k + 1] = nn.y;
                normal[3 * k + 2] = nn.z;
                depth[k] = d;
            }
            if (stats) {
                stats[k * 4 + 0] = ms.peakFaces;
                stats[k * 4 + 1] = ms.peakEdges;
                stats[k * 4 + 2] = ms.iterations;
                stats[k * 4 + 3] = uint32_t(std::min<uint64_t>(ms.peakMultiplicity, 0xffffffffu));
            }
        }
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// Same contract as ref_gjk_epa.
int orc_gjk_epa(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                uint32_t gjkIterations, uint32_t epaIterations, float tolerance, int32_t* status,
                float* normal, float* depth, float* simplex)
{
    try {
        View<float> v(c);
        for (uint64_t k = 0; k < n; ++k) {
            Shape<float> sa = v.shape(pairs[k].first), sb = v.shape(pairs[k].second);
            Pose<float> ta = v.pose(pairs[k].first), tb = v.pose(pairs[k].second);
            np::Support<float> fa{&sa, &ta}, fb{&sb, &tb};
            np::Simplex<float> s;
            bool hitG = np::gjk(fa, fb, s, gjkIterations);
            if (simplex) {
                for (int p = 0; p < 4; ++p) {
                    simplex[k * 13 + p * 3 + 0] = s.p[p].x;
                    simplex[k * 13 + p * 3 + 1] = s.p[p].y;
                    simplex[k * 13 + p * 3 + 2] = s.p[p].z;
                }
                simplex[k * 13 + 12] = float(s.size);
            }
            normal[3 * k] = normal[3 * k + 1] = normal[3 * k + 2] = 0.0f;
            depth[k] = 0.0f;
            if (!hitG) {
                status[k] = 0;
                continue;
            }
            V3<float> nn;
            float d;
            NarrowOptions opt = makeOptions(gjkIterations, epaIterations, tolerance);
            opt.epaIterations = epaIterations;
            const int st = np::epa(fa, fb, s, opt, nn, d);
            status[k] = st == 1 ? 1 : (st == 0 ? 2 : (st == -1 ? 3 : 4));
            normal[3 * k] = nn.x;
            normal[3 * k + 1] = nn.y;
            normal[3 * k + 2] = nn.z;
            depth[k] = d;
        }
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// World-level collision section (physics_world.cpp:539-739) with the same
// contract as ref_world_collide, hulls included: body pose (identity
// rotation) composed with the collider local pose, broad phase, mapping,
// same-body filter (one body per collider here, so nothing is removed),
// narrow phase, anchors to body-local, material combination, compaction in
// candidate order.
static int worldCollide(const physgpu_colliders* c, const float* bodyPosition,
                        const uint32_t* colliderBody,
                      const float* localPose, const float* friction, const float* restitution,
                      int algorithm, uint32_t bpWorkers, uint32_t npWorkers, uint32_t warmup,
                      uint32_t steps, physgpu_manifold* out, uint64_t capacity, uint64_t* count,
                      double* timing, uint64_t* npStats)
{
    try {
        using Clock = std::chrono::steady_clock;
        auto ms = [](Clock::time_point a, Clock::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        const uint32_t n = c->count;
        View<float> v(c);
        std::vector<Shape<float>> shapes(n);
        std::vector<Pose<float>> local(n), world(n);
        for (uint32_t i = 0; i < n; ++i) {
            shapes[i] = v.shape(i);
            if (localPose)
                local[i] = {{localPose[i], localPose[n + i], localPose[2 * n + i]},
                            {localPose[3 * n + i], localPose[4 * n + i], localPose[5 * n + i],
                             localPose[6 * n + i]}};
            else
                local[i] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 1.0f}};
        }
        NarrowOptions opt = makeOptions(0, 0, 0.0f);
        std::vector<physgpu_aabb> bounds(n);
        std::vector<Pair> pairs;
        std::vector<physgpu_manifold> results;
        std::vector<uint8_t> hits;
        std::vector<double> tb, tn, ta;
        uint64_t stats6[6] = {0, 0, 0, 0, 0, 0};
        std::vector<physgpu_manifold> contacts;
        for (uint32_t s = 0; s < warmup + steps; ++s) {
            auto t0 = Clock::now();
            for (uint32_t i = 0; i < n; ++i) {
                Pose<float> body{{bodyPosition[i], bodyPosition[n + i], bodyPosition[2 * n + i]},
                                 {0.0f, 0.0f, 0.0f, 1.0f}};
                world[i] = {add(body.p, rotate(body.q, local[i].p)), qmul(body.q, local[i].q)};
                Box6<float> b = aabbOf(shapes[i], world[i]);
                bounds[i] = {{b.mn.x, b.mn.y, b.mn.z}, {b.mx.x, b.mx.y, b.mx.z}};
            }
            auto t1 = Clock::now();
            physgpu_broadphase_stats st;
            findPairs(bounds.data(), n, algorithm, bpWorkers, pairs, st);
            // Pairs of colliders on one body are erased, order kept
            // (physics_world.cpp:630-633).
            if (colliderBody)
                pairs.erase(std::remove_if(pairs.begin(), pairs.end(),
                                           [&](const Pair& q) {
                                               return colliderBody[q.first] ==
                                                      colliderBody[q.second];
                                           }),
                            pairs.end());
            auto t2 = Clock::now();
            results.resize(pairs.size());
            hits.assign(pairs.size(), 0);
            auto one = [&](size_t k) {
                uint32_t a = pairs[k].first, b = pairs[k].second;
                Contact<float> ct;
                bool h = generateContact(shapes[a], world[a], shapes[b], world[b], opt, ct);
                hits[k] = h;
                if (!h)
                    return;
                for (uint32_t p = 0; p < ct.count; ++p) {
                    ct.anchorA[p] = add(local[a].p, rotate(local[a].q, ct.anchorA[p]));
                    ct.anchorB[p] = add(local[b].p, rotate(local[b].q, ct.anchorB[p]));
                }
                physgpu_manifold& m = results[k];
                std::memset(&m, 0, sizeof(m));
                m.first = a;
                m.second = b;
                m.pair_index = uint32_t(k);
                m.shape_pair = pairClass(shapes[a].type, shapes[b].type);
                store(ct, &m);
                float ra = restitution ? restitution[a] : 0.0f;
                float rb = restitution ? restitution[b] : 0.0f;
                float fa = friction ? friction[a] : 0.0f;
                float fb = friction ? friction[b] : 0.0f;
                m.restitution = maxOf(ra, rb);
                m.friction = std::sqrt(maxOf(fa, 0.0f) * maxOf(fb, 0.0f));
            };
            Pool::run(pairs.size() >= 4096 ? npWorkers : 1, pairs.size(), one);
            contacts.clear();
            std::fill(stats6, stats6 + 6, 0);
            for (size_t k = 0; k < pairs.size(); ++k) {
                uint32_t cls = pairClass(shapes[pairs[k].first].type, shapes[pairs[k].second].type);
                if (cls < 3)
                    ++stats6[cls];
                if (!hits[k])
                    continue;
                contacts.push_back(results[k]);
                if (cls < 3)
                    ++stats6[3 + cls];
            }
            auto t3 = Clock::now();
            if (s >= warmup) {
                ta.push_back(ms(t0, t1));
                tb.push_back(ms(t1, t2));
                tn.push_back(ms(t2, t3));
            }
        }
        auto median = [](std::vector<double> x) {
            if (x.empty())
                return 0.0;
            std::sort(x.begin(), x.end());
            return x[x.size() / 2];
        };
        if (timing) {
            timing[0] = median(tb);
            timing[1] = median(tn);
            timing[2] = median(ta);
            timing[3] = timing[0] + timing[1] + timing[2];
            timing[4] = double(pairs.size());
            // possiblePairs (physics_world.cpp:571-577): the earlier
            // colliders on other bodies, summed over colliders.
            uint64_t possible = 0;
            std::vector<uint64_t> perBody(colliderBody ? n : 0);
            for (uint32_t i = 0; i < n; ++i) {
                possible += i;
                if (colliderBody)
                    possible -= perBody[colliderBody[i]]++;
            }
            timing[5] = double(possible);
        }
        if (npStats)
            std::copy(stats6, stats6 + 6, npStats);
        *count = contacts.size();
        if (out)
            for (size_t k = 0; k < contacts.size() && k < capacity; ++k)
                out[k] = contacts[k];
        return contacts.size() > capacity ? PHYSGPU_CAPACITY_EXCEEDED : PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// One body per collider (bodyPosition, friction and restitution per collider).
int orc_world_collide(const physgpu_colliders* c, const float* bodyPosition,
                      const float* localPose, const float* friction, const float* restitution,
                      int algorithm, uint32_t bpWorkers, uint32_t npWorkers, uint32_t warmup,
                      uint32_t steps, physgpu_manifold* out, uint64_t capacity, uint64_t* count,
                      double* timing, uint64_t* npStats)
{
    return worldCollide(c, bodyPosition, nullptr, localPose, friction, restitution, algorithm,
                        bpWorkers, npWorkers, warmup, steps, out, capacity, count, timing,
                        npStats);
}

// Colliders sharing bodies: colliderBody[i] is the body of collider i (index
// < n); bodyPosition, friction and restitution are given per collider (its
// body's values), as the HIP library's physgpu_set_world takes them.
int orc_world_collide_bodies(const physgpu_colliders* c, const float* bodyPosition,
                             const uint32_t* colliderBody, const float* localPose,
                             const float* friction, const float* restitution, int algorithm,
                             uint32_t bpWorkers, uint32_t npWorkers, uint32_t warmup,
                             uint32_t steps, physgpu_manifold* out, uint64_t capacity,
                             uint64_t* count, double* timing, uint64_t* npStats)
{
    for (uint32_t i = 0; i < c->count; ++i)
        if (colliderBody[i] >= c->count)
            return fail(std::invalid_argument("body index out of range"));
    return worldCollide(c, bodyPosition, colliderBody, localPose, friction, restitution,
                        algorithm, bpWorkers, npWorkers, warmup, steps, out, capacity, count,
                        timing, npStats);
}

// Pairs are the manifolds' (first, second).  status[k]: 1 differentiated,
// 0 no contact in FP64, -1 hull pair (not differentiated).  grad: 14 x n
// planes (A's px py pz qx qy qz qw, then B's).
int orc_contact_adjoint_f64(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                            const float* gn, const float* gd, const float* gp, uint32_t workers,
                            double* grad, int32_t* status)
{
    try {
        View<double> v(c);
        auto one = [&](size_t k) {
            const uint32_t a = pairs[k].first, b = pairs[k].second;
            for (int q = 0; q < 14; ++q)
                grad[q * n + k] = 0.0;
            if (sphereHull(v, a, b)) {
                status[k] = -1;
                return;
            }
            double pa[7], pb[7];
            posesOf(c, a, pa);
            posesOf(c, b, pb);
            Pose<Dual> ta = normalisedPoseSeeded(pa, 0), tb = normalisedPoseSeeded(pb, 7);
            Contact<Dual> ct;
            if (!contactTyped(v, a, b, ta, tb, ct, nullptr)) {
                status[k] = 0;
                return;
            }
            const Dual phi = Upstream(gn, gd, gp, n, k).phi(ct);
            for (int q = 0; q < 14; ++q)
                grad[q * n + k] = phi.d[q];
            status[k] = 1;
        };
        Pool::run(workers, n, one);
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// Central differences (step h) of the linearised loss in FP64, each pose
// component perturbed in turn (quaternions perturbed, then normalised).
// stable[k] = 1 when every perturbed evaluation keeps the contact and all its
// discrete choices.
int orc_contact_fd_f64(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                       const float* gn, const float* gd, const float* gp, double h,
                       uint32_t workers, double* grad, int32_t* stable)
{
    try {
        View<double> v(c);
        auto one = [&](size_t k) {
            const uint32_t a = pairs[k].first, b = pairs[k].second;
            stable[k] = 0;
            for (int q = 0; q < 14; ++q)
                grad[q * n + k] = 0.0;
            if (sphereHull(v, a, b))
                return;
            const Shape<double> sa = v.shape(a), sb = v.shape(b);
            const bool hull = sa.type == kHull || sb.type == kHull;
            double x[14];
            posesOf(c, a, x);
            posesOf(c, b, x + 7);
            const Upstream up(gn, gd, gp, n, k);
            Contact<double> c0;
            Choice ch0;
            if (!contactF64(v, a, b, x, x + 7, c0, &ch0))
                return;
            bool same = true;
            for (int q = 0; q < 14; ++q) {
                double f[2] = {0.0, 0.0};
                for (int s = 0; s < 2; ++s) {
                    double y[14];
                    std::copy(x, x + 14, y);
                    y[q] += s ? -h : h;
                    Contact<double> cs;
                    Choice ch;
                    if (!contactF64(v, a, b, y, y + 7, cs, &ch) || !(ch == ch0) ||
                        cs.count != c0.count)
                        same = false;
                    // Hull pairs: the EPA face is unchanged when normal and
                    // depth move continuously; the contact point (midpoint of
                    // the support points) jumps when a support is a tied face
                    // or edge, which only matters when its seed is non-zero.
                    if (hull && same) {
                        const double dn = std::max({std::abs(cs.normal.x - c0.normal.x),
                                                    std::abs(cs.normal.y - c0.normal.y),
                                                    std::abs(cs.normal.z - c0.normal.z)});
                        const double dd = std::abs(cs.depth[0] - c0.depth[0]);
                        const double dw = std::max({std::abs(cs.world[0].x - c0.world[0].x),
                                                    std::abs(cs.world[0].y - c0.world[0].y),
                                                    std::abs(cs.world[0].z - c0.world[0].z)});
                        const bool pointSeed = up.p[0][0] != 0.0 || up.p[0][1] != 0.0 ||
                                               up.p[0][2] != 0.0;
                        same = dn < 1e3 * h && dd < 1e3 * h && (!pointSeed || dw < 1e3 * h);
                    }
                    f[s] = up.phi(cs);
                }
                grad[q * n + k] = (f[0] - f[1]) / (2.0 * h);
            }
            stable[k] = same ? 1 : 0;
        };
        Pool::run(workers, n, one);
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// Discrete choices of each pair's FP64 contact, 12 ints per pair: hit, sel,
// flip, fallback, count, branch, face, clamped, support[0], support[1],
// and the number of zero-clamped clipping depths, kept points from B's face.
int orc_contact_choices(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                        int32_t* out)
{
    try {
        View<double> v(c);
        for (uint64_t k = 0; k < n; ++k) {
            double pa[7], pb[7];
            posesOf(c, pairs[k].first, pa);
            posesOf(c, pairs[k].second, pb);
            Contact<double> ct{};
            Choice ch;
            const bool hit = contactF64(v, pairs[k].first, pairs[k].second, pa, pb, ct, &ch);
            int zeros = 0, fromB = 0;
            for (int j = 0; j < ch.count; ++j) {
                zeros += ch.zero[j];
                fromB += ch.side[j];
            }
            const int32_t row[12] = {hit, ch.sel, ch.flip, ch.fallback, ch.count, ch.branch,
                                     ch.face, ch.clamped, ch.support[0], ch.support[1], zeros,
                                     fromB};
            std::copy(row, row + 12, out + 12 * k);
        }
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

// FP64 normal (3 x n), depths (4 x n), world points (12 x n) and point count
// of each pair (0 = no contact or hull pair), orientation q/|q|.
int orc_contact_geometry_f64(const physgpu_colliders* c, const physgpu_pair* pairs, uint64_t n,
                             double* normal, double* depth, double* point, int32_t* count)
{
    try {
        View<double> v(c);
        for (uint64_t k = 0; k < n; ++k) {
            double pa[7], pb[7];
            posesOf(c, pairs[k].first, pa);
            posesOf(c, pairs[k].second, pb);
            Contact<double> ct{};
            const bool hit = contactF64(v, pairs[k].first, pairs[k].second, pa, pb, ct, nullptr);
            count[k] = hit ? int32_t(ct.count) : 0;
            const double nv[3] = {ct.normal.x, ct.normal.y, ct.normal.z};
            for (int a = 0; a < 3; ++a)
                normal[a * n + k] = hit ? nv[a] : 0.0;
            for (int j = 0; j < 4; ++j) {
                const bool used = hit && uint32_t(j) < ct.count;
                depth[j * n + k] = used ? ct.depth[j] : 0.0;
                point[(3 * j) * n + k] = used ? ct.world[j].x : 0.0;
                point[(3 * j + 1) * n + k] = used ? ct.world[j].y : 0.0;
                point[(3 * j + 2) * n + k] = used ? ct.world[j].z : 0.0;
            }
        }
        return PHYSGPU_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

} // extern "C"
