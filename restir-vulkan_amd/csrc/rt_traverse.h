// rt_traverse.h — AABB-tree traversal on gfx950 over the reference's node layout
// (aabbTree.glsl:1-13: both child boxes stored in the parent, child < 0 = leaf ~triIndex).
//
// Any-hit shadow traversal follows softwareRaytracing.glsl:42-85: depth-first from node 0, the
// left child pushed before the right one (so the right subtree is visited first). The per-lane
// stack lives in LDS, one 4-byte slot per lane per level, laid out [level][lane] so a wave's 64
// lanes hit 64 consecutive dwords. The node about to be visited is kept in a register instead
// of being pushed and popped at once, which leaves the visiting order unchanged.
//
// Two loop disciplines:
//   kRefLoop = true : the reference's exact loop: triangle tests become visible only every 8
//                     node visits (geomTestInterval) or when the stack runs empty, so the
//                     visit count equals the reference's loop-iteration count.
//   kRefLoop = false: the production loop: a ray stops at its first triangle hit.
// Both return the same occlusion result.
#pragma once

#include "rt_math.h"

namespace rt {

struct alignas(16) Node {
    float4 lmin, lmax, rmin, rmax;
    int32_t left, right, pad0, pad1;
};
static_assert(sizeof(Node) == 80, "AabbTreeNode is 80 bytes");

// rayAabIntersection, softwareRaytracing.glsl:9-14 (segment parametrised on [0, 1]).
__device__ __forceinline__ bool ray_box_div(f3 o, f3 d, float4 mn, float4 mx) {
    const float ax = (mn.x - o.x) / d.x, ay = (mn.y - o.y) / d.y, az = (mn.z - o.z) / d.z;
    const float bx = (mx.x - o.x) / d.x, by = (mx.y - o.y) / d.y, bz = (mx.z - o.z) / d.z;
    const float rmin = gmax(gmin(ax, bx), gmax(gmin(ay, by), gmin(az, bz)));
    const float rmax = gmin(gmax(ax, bx), gmin(gmax(ay, by), gmax(az, bz)));
    return rmin < 1.0f && rmax >= rmin && rmax > 0.0f;
}

// A segment with its per-component reciprocal direction (1/d correctly rounded, once per ray).
struct Seg {
    f3 o, d, inv;
};
__device__ __forceinline__ Seg make_seg(f3 o, f3 d) {
    return Seg{o, d, f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z}};
}

// Same decision as ray_box_div, without its six IEEE divisions in the common case.
// Each slab distance is formed as RN(RN(b - o) * RN(1/d)); the reference's RN(RN(b - o) / d)
// differs from it by at most 3 * 2^-24 relative (three roundings of a product of the same exact
// numerator), and min/max/max3/min3 carry that bound to rmin and rmax. If the three comparisons
// come out the same at both ends of the widened intervals [v(1-8u), v(1+8u)] (u = 2^-24, plus an
// absolute 1e-30 for results near zero), the reference's answer is known; otherwise — and for
// any non-finite or huge distance — the division form is evaluated. The slack is four times the
// bound, so the fallback is taken only for segments grazing a box face within ~1e-6.
__device__ __forceinline__ bool ray_box(const Seg &s, float4 mn, float4 mx) {
    const float ax = (mn.x - s.o.x) * s.inv.x, ay = (mn.y - s.o.y) * s.inv.y, az = (mn.z - s.o.z) * s.inv.z;
    const float bx = (mx.x - s.o.x) * s.inv.x, by = (mx.y - s.o.y) * s.inv.y, bz = (mx.z - s.o.z) * s.inv.z;
    const float rmin = gmax(gmin(ax, bx), gmax(gmin(ay, by), gmin(az, bz)));
    const float rmax = gmin(gmax(ax, bx), gmin(gmax(ay, by), gmax(az, bz)));
    // a comparison is false for NaN, so this also rejects NaN and infinities
    const bool finite = fabsf(ax) < 1e30f && fabsf(bx) < 1e30f && fabsf(ay) < 1e30f &&
                        fabsf(by) < 1e30f && fabsf(az) < 1e30f && fabsf(bz) < 1e30f;
    constexpr float kRel = 8.0f * 5.9604645e-8f, kAbs = 1e-30f;
    const float emin = fabsf(rmin) * kRel + kAbs, emax = fabsf(rmax) * kRel + kAbs;
    const float minLo = rmin - emin, minHi = rmin + emin;
    const float maxLo = rmax - emax, maxHi = rmax + emax;
    const bool surelyIn = minHi < 1.0f && maxLo >= minHi && maxLo > 0.0f;
    const bool surelyOut = minLo >= 1.0f || maxHi < minLo || maxHi <= 0.0f;
    if (finite && (surelyIn || surelyOut)) return surelyIn;
    return ray_box_div(s.o, s.d, mn, mx);
}

// rayTriangleIntersection, softwareRaytracing.glsl:15-37 (Moller-Trumbore, t in (0, 1)).
__device__ __forceinline__ bool ray_tri(const float4 *__restrict__ tris, int32_t t, f3 o, f3 d) {
    const float4 a = tris[3 * t + 0], b = tris[3 * t + 1], c = tris[3 * t + 2];
    const f3 p1 = xyz(a);
    const f3 e1 = xyz(b) - p1;
    const f3 e2 = xyz(c) - p1;
    const f3 p = cross(d, e2);
    float f = 1.0f / dot(e1, p);
    const f3 s = o - p1;
    const float bx = f * dot(s, p);
    if (bx < 0.0f || bx > 1.0f) return false;
    const f3 q = cross(s, e1);
    const float by = f * dot(d, q);
    if (by < 0.0f || by + bx > 1.0f) return false;
    f *= dot(e2, q);
    return f > 0.0f && f < 1.0f;
}

__device__ __forceinline__ Node load_node(const Node *__restrict__ nodes, int32_t i) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + i);
    Node n;
    n.lmin = p[0];
    n.lmax = p[1];
    n.rmin = p[2];
    n.rmax = p[3];
    const int4 ch = reinterpret_cast<const int4 *>(p)[4];
    n.left = ch.x;
    n.right = ch.y;
    n.pad0 = n.pad1 = 0;
    return n;
}

struct TraceStats {
    uint32_t visits, tests;
};

// Returns true if the segment o + t*d, t in (0,1), hits a triangle. `stk` points at this lane's
// slot 0 of the LDS stack; consecutive levels are `stride` dwords apart.
template <int kStack, bool kRefLoop>
__device__ __forceinline__ bool occluded(const Node *__restrict__ nodes,
                                         const float4 *__restrict__ tris, f3 o, f3 d,
                                         int32_t *stk, int stride, TraceStats &st) {
    const Seg s = make_seg(o, d);
    int top = 0;
    int32_t cur = 0;
    int counter = 0;
    bool hit = false;
    uint32_t nv = 0, nt = 0;
    for (;;) {
        if (cur < 0) {
            if (top == 0) break;
            --top;
            cur = stk[top * stride];
        }
        const Node nd = load_node(nodes, cur);
        ++nv;
        const bool li = ray_box(s, nd.lmin, nd.lmax);
        const bool ri = ray_box(s, nd.rmin, nd.rmax);
        int32_t next = -1;
        if (li) {
            if (nd.left < 0) {
                if (!hit) {
                    ++nt;
                    if (ray_tri(tris, ~nd.left, o, d)) {
                        hit = true;
                        if (!kRefLoop) break;
                    }
                }
            } else {
                next = nd.left;
            }
        }
        if (ri) {
            if (nd.right < 0) {
                if (!hit) {
                    ++nt;
                    if (ray_tri(tris, ~nd.right, o, d)) {
                        hit = true;
                        if (!kRefLoop) break;
                    }
                }
            } else {
                if (next >= 0 && top < kStack) {
                    stk[top * stride] = next;
                    ++top;
                }
                next = nd.right;
            }
        }
        cur = next;
        if (kRefLoop) {
            if (++counter == 8) {
                if (hit) break;
                counter = 0;
            }
        }
    }
    st.visits += nv;
    st.tests += nt;
    return hit;
}

// Step-wise form of `occluded` for schedules that interleave rays on one lane: the state holds one
// ray's traversal, and each step visits exactly one node (the same visiting order, node count and
// result as `occluded`). A step returns true when the ray is finished; `hit` then has the answer.
struct Walk {
    Seg s;
    int32_t cur;
    int top;
    int counter;
    bool hit;
    uint32_t nv, nt;
};

__device__ __forceinline__ void walk_begin(Walk &w, f3 o, f3 d) {
    w.s = make_seg(o, d);
    w.cur = 0;
    w.top = 0;
    w.counter = 0;
    w.hit = false;
    w.nv = w.nt = 0;
}

template <int kStack, bool kRefLoop>
__device__ __forceinline__ bool walk_step(Walk &w, const Node *__restrict__ nodes,
                                          const float4 *__restrict__ tris, int32_t *stk, int stride) {
    const Node nd = load_node(nodes, w.cur);
    ++w.nv;
    const bool li = ray_box(w.s, nd.lmin, nd.lmax);
    const bool ri = ray_box(w.s, nd.rmin, nd.rmax);
    int32_t next = -1;
    bool stop = false;
    if (li) {
        if (nd.left < 0) {
            if (!w.hit) {
                ++w.nt;
                if (ray_tri(tris, ~nd.left, w.s.o, w.s.d)) {
                    w.hit = true;
                    stop = !kRefLoop;
                }
            }
        } else {
            next = nd.left;
        }
    }
    if (ri && !stop) {
        if (nd.right < 0) {
            if (!w.hit) {
                ++w.nt;
                if (ray_tri(tris, ~nd.right, w.s.o, w.s.d)) {
                    w.hit = true;
                    stop = !kRefLoop;
                }
            }
        } else {
            if (next >= 0 && w.top < kStack) {
                stk[w.top * stride] = next;
                ++w.top;
            }
            next = nd.right;
        }
    }
    if (stop) return true;
    if (kRefLoop) {
        if (++w.counter == 8) {
            if (w.hit) return true;
            w.counter = 0;
        }
    }
    if (next < 0) {
        if (w.top == 0) return true;
        --w.top;
        next = stk[w.top * stride];
    }
    w.cur = next;
    return false;
}

// testVisibility, visibilityTest.glsl:1-30: shrink the segment p1 -> p2 by tMin = 0.001 at both
// ends. Returns the ray origin and the (unnormalised) direction.
__device__ __forceinline__ void shadow_segment(f3 p1, f3 p2, f3 &o, f3 &d) {
    const float tMin = 0.001f;
    const f3 dir = p2 - p1;
    const f3 off = tMin * normalize(dir);
    o = p1 + off;
    d = dir - 2.0f * off;
}

}  // namespace rt
