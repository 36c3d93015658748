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
// The decision is used only when all six distances are finite, hence not NaN; there fminf/fmaxf
// (one v_min/v_max or v_min3/v_max3 each) select the same values as the reference's a<b?a:b
// forms, up to the sign of a zero, which no comparison below can see (emin/emax take |.|).
__device__ __forceinline__ bool ray_box(const Seg &s, float4 mn, float4 mx) {
    const float ax = (mn.x - s.o.x) * s.inv.x, ay = (mn.y - s.o.y) * s.inv.y, az = (mn.z - s.o.z) * s.inv.z;
    const float bx = (mx.x - s.o.x) * s.inv.x, by = (mx.y - s.o.y) * s.inv.y, bz = (mx.z - s.o.z) * s.inv.z;
    const float rmin = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    const float rmax = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
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

// ---------------------------------------------------------------- 4-entry walk of the tree --
// The same tree, two levels per visit. WNode x (one per internal node x of the reference tree,
// same index) holds, for each side of x, either that child itself when it is a triangle leaf, or
// the child's two children: up to four entries, each with the box the reference tree stores for
// it. Both children's boxes sit in their parent, so the entries are copied, not recomputed.
//
// Shadow rays only ask whether some triangle on the segment is hit, and the reference tests a
// triangle exactly when every box on its path from the root passes. A visit tests its entries'
// boxes but not the boxes of the internal children between them and the visited node. That
// skips no triangle and adds none when each skipped box passes whenever a box inside it passes:
// ray_box decides as the division form RN(RN(b - o) / d) does, each slab distance is monotone
// in the plane b for a fixed ray, and with finite planes and o and a finite nonzero d no distance
// is NaN, so a box nested in another yields an interval [rmin, rmax] inside the other's and the
// pass condition carries over. For a ray with a zero or non-finite component of d or o (a slab
// distance 0/0 is NaN, and the reference's a<b?a:b chains then drop that slab), and for a tree
// whose boxes are not known to be finite and nested, the skipped boxes are tested as well, from
// the reference node of the visited node: every box on every path is then tested, as in the
// binary walk.
// The answer is the reference's as long as no push is dropped: each visit pushes at most one item
// and each visit is one or two levels below the one that pushed it, so at most (level of the
// current node) items wait, no more than the depth - 1 of the binary walk.
//
// Per side of a visited node x with internal child c: if both of c's entries are internal and
// pass, the side yields one item, c itself (its visit tests the entries below both); if one
// passes, the side yields that entry; leaves that pass are tested at once.
struct alignas(16) WNode {
    float4 lox, loy, loz;  // low corner of entry k in component k
    float4 hix, hiy, hiz;  // high corner
    int4 id;               // entry: >= 0 internal node, < 0 ~triangle, kNoEntry: none
    int4 grp;              // .x / .y: the left / right child when it is internal, else -1;
                           // .z: bit k when entry k exists, bit 4 + k when it is a triangle
};
static_assert(sizeof(WNode) == 128, "one 128-byte line per node");
constexpr int32_t kNoEntry = INT32_MIN;

// A segment for the 4-entry walk. `full`: test the skipped boxes too (see above). `quick`: the
// tree's boxes are finite, ordered and nested, o and d are regular, and every slab distance of
// every box of the tree stays below 1e30 in magnitude, because (max |box coordinate| + max |o|) *
// max |1/d| < 1e30; box_quick then forms each slab distance with one FMA and needs no finiteness
// test. ox, oy, oz: byte offsets of the entries' near planes in a WNode row pair (0 when that
// component of d is positive, so the low corner is met first, else 48).
struct WSeg {
    Seg s;
    f3 noi;     // -RN(o * inv), per component
    float eq;   // slack of the quick decision: 32u + 4u max |o * inv|, u = 2^-24
    uint32_t ox, oy, oz;
    bool full, quick;
};
__device__ __forceinline__ WSeg make_wseg(f3 o, f3 d, bool treeNested, float maxCoord) {
    constexpr float kInf = __builtin_huge_valf();
    // false for NaN as well as for zero or infinite components
    const bool regular = fabsf(d.x) > 0.0f && fabsf(d.y) > 0.0f && fabsf(d.z) > 0.0f &&
                         fabsf(d.x) < kInf && fabsf(d.y) < kInf && fabsf(d.z) < kInf &&
                         fabsf(o.x) < kInf && fabsf(o.y) < kInf && fabsf(o.z) < kInf;
    WSeg w;
    w.s = make_seg(o, d);
    const f3 inv = w.s.inv;
    w.noi = f3{-(o.x * inv.x), -(o.y * inv.y), -(o.z * inv.z)};
    const float invMax = fmaxf(fmaxf(fabsf(inv.x), fabsf(inv.y)), fabsf(inv.z));
    const float invMin = fminf(fminf(fabsf(inv.x), fabsf(inv.y)), fabsf(inv.z));
    const float oMax = fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z));
    w.full = !(treeNested && regular);
    // comparisons are false for NaN: a NaN maxCoord or an infinite 1/d leaves quick off
    w.quick = !w.full && invMin > 1e-30f && (maxCoord + oMax) * invMax < 1e30f;
    const float oiMax = fmaxf(fmaxf(fabsf(w.noi.x), fabsf(w.noi.y)), fabsf(w.noi.z));
    w.eq = fmaf(oiMax, 2.3841858e-7f, 1.9073486e-6f);
    w.ox = d.x > 0.0f ? 0u : 48u;
    w.oy = d.y > 0.0f ? 0u : 48u;
    w.oz = d.z > 0.0f ? 0u : 48u;
    return w;
}

// ray_box for a quick WSeg, given each axis's near and far plane (the low and high corner in the
// order the segment meets them). Each slab distance is t = RN(b * inv - oi) with one FMA
// (oi = RN(o * inv)); with inv = (1/d)(1 + d1) and oi = o inv (1 + d2), t differs from the exact
// (b - o)/d by at most 2.01u |t| + 1.01u |o inv|, so from the reference's RN(RN(b - o)/d) by at
// most E(t) = 5.03u |t| + 1.02u A (A = max |o inv|). t is monotone in b for a fixed ray, in both
// forms, so near <= far per axis, rmin = max3 of the near distances and rmax = min3 of the far
// ones are the reference's max-of-mins and min-of-maxes, and the max and min of values each within
// E of its counterpart stay within E of the counterpart's max and min (E grows with |t|, and t +-
// E(t) is monotone). The decision rmin < 1, rmax >= rmin, rmax > 0 is then the reference's as soon
// as |rmin - 1|, |rmax| and |rmax - rmin| are all at least eq = 32u + 4uA: each of the three
// differences is then larger than the error of its two sides (for |t| <= 2 the errors are below
// 10.1u + 1.02uA per side; for larger |t| the distance to the threshold grows with |t| while
// 5.03u |t| does not catch up), except rmax - rmin when |rmin| or |rmax| exceeds 2, and there
// either rmin > 1 or rmax < 0 already decides both forms, or rmax - rmin > 1 in both. Closer
// cases take the division form, which gives the same decision with near and far as with the low
// and high corners (its per-axis min and max only swap equal operands). All distances are finite
// by the `quick` condition, so fminf/fmaxf see no NaN.
__device__ __forceinline__ bool box_quick(const WSeg &w, float nx, float ny, float nz, float fx, float fy, float fz) {
    const f3 inv = w.s.inv, noi = w.noi;
    const float rmin = fmaxf(fmaxf(fmaf(nx, inv.x, noi.x), fmaf(ny, inv.y, noi.y)), fmaf(nz, inv.z, noi.z));
    const float rmax = fminf(fminf(fmaf(fx, inv.x, noi.x), fmaf(fy, inv.y, noi.y)), fmaf(fz, inv.z, noi.z));
    const float t1 = rmin - 1.0f, gap = rmax - rmin;
    if (fminf(fminf(fabsf(t1), fabsf(rmax)), fabsf(gap)) >= w.eq) return t1 < 0.0f && gap >= 0.0f && rmax > 0.0f;
    return ray_box_div(w.s.o, w.s.d, make_float4(nx, ny, nz, 0.0f), make_float4(fx, fy, fz, 0.0f));
}

__device__ __forceinline__ float4 corner(float4 x, float4 y, float4 z, int k) {
    return k == 0 ? make_float4(x.x, y.x, z.x, 0.0f)
         : k == 1 ? make_float4(x.y, y.y, z.y, 0.0f)
         : k == 2 ? make_float4(x.z, y.z, z.z, 0.0f)
                  : make_float4(x.w, y.w, z.w, 0.0f);
}

// The entries of WNode `cur` that the segment passes: bit k for entry k when it exists, its box
// passes and (under ws.full) so does the box of its side. `id` / `grp` get the node's entry and
// group words. kQuick: every active lane's ray is quick; each axis's near and far rows are then
// loaded by the ray's direction, so the four box tests need no per-axis min and max.
template <bool kQuick>
__device__ __forceinline__ uint32_t wide_pass(const WSeg &ws, const WNode *__restrict__ wide,
                                              const Node *__restrict__ nodes, int32_t cur, int4 &id, int4 &grp) {
    uint32_t pass;
    if (kQuick) {
        // row pairs (low, high) of x, y, z at byte offsets (0, 48), (16, 64), (32, 80) of the node:
        // 32-bit offsets from the uniform base, so each load takes one offset register
        const char *base = reinterpret_cast<const char *>(wide);
        const uint32_t row = static_cast<uint32_t>(cur) << 7;
        const float4 nx = *reinterpret_cast<const float4 *>(base + (row + ws.ox));
        const float4 fx = *reinterpret_cast<const float4 *>(base + (row + (ws.ox ^ 48u)));
        const float4 ny = *reinterpret_cast<const float4 *>(base + (row + ws.oy) + 16);
        const float4 fy = *reinterpret_cast<const float4 *>(base + (row + (ws.oy ^ 48u)) + 16);
        const float4 nz = *reinterpret_cast<const float4 *>(base + (row + ws.oz) + 32);
        const float4 fz = *reinterpret_cast<const float4 *>(base + (row + (ws.oz ^ 48u)) + 32);
        id = *reinterpret_cast<const int4 *>(base + row + 96);
        grp = *reinterpret_cast<const int4 *>(base + row + 112);
        // all four boxes are tested (an empty entry holds a zero box), without a branch per entry
        const bool b0 = box_quick(ws, nx.x, ny.x, nz.x, fx.x, fy.x, fz.x);
        const bool b1 = box_quick(ws, nx.y, ny.y, nz.y, fx.y, fy.y, fz.y);
        const bool b2 = box_quick(ws, nx.z, ny.z, nz.z, fx.z, fy.z, fz.z);
        const bool b3 = box_quick(ws, nx.w, ny.w, nz.w, fx.w, fy.w, fz.w);
        pass = (b0 ? 1u : 0u) | (b1 ? 2u : 0u) | (b2 ? 4u : 0u) | (b3 ? 8u : 0u);
    } else {
        const float4 *p = reinterpret_cast<const float4 *>(wide + cur);
        id = reinterpret_cast<const int4 *>(p)[6];
        grp = reinterpret_cast<const int4 *>(p)[7];
        const float4 lx = p[0], ly = p[1], lz = p[2], hx = p[3], hy = p[4], hz = p[5];
        bool sideL = true, sideR = true;
        if (ws.full) {
            const Node nd = load_node(nodes, cur);
            sideL = ray_box(ws.s, nd.lmin, nd.lmax);
            sideR = ray_box(ws.s, nd.rmin, nd.rmax);
        }
        const bool b0 = ray_box(ws.s, corner(lx, ly, lz, 0), corner(hx, hy, hz, 0));
        const bool b1 = ray_box(ws.s, corner(lx, ly, lz, 1), corner(hx, hy, hz, 1));
        const bool b2 = ray_box(ws.s, corner(lx, ly, lz, 2), corner(hx, hy, hz, 2));
        const bool b3 = ray_box(ws.s, corner(lx, ly, lz, 3), corner(hx, hy, hz, 3));
        const uint32_t sides = (sideL ? 3u : 0u) | (sideR ? 12u : 0u);
        pass = ((b0 ? 1u : 0u) | (b1 ? 2u : 0u) | (b2 ? 4u : 0u) | (b3 ? 8u : 0u)) & sides;
    }
    return pass & static_cast<uint32_t>(grp.z);
}

// The triangle entries among the passed entries `h`.
__device__ __forceinline__ uint32_t wide_leaves(uint32_t h, int4 grp) {
    return h & (static_cast<uint32_t>(grp.z) >> 4);
}

__device__ __forceinline__ int32_t wide_entry(int4 id, uint32_t k) {
    return k == 0 ? id.x : (k == 1 ? id.y : (k == 2 ? id.z : id.w));
}

// The node visited after the passed entries `h`: per side, its internal child itself when both of
// that child's entries passed, else the one entry that passed; the left side's item is pushed
// when both sides yield one; with neither, the stack's top (-1 when the stack is empty).
template <int kStack>
__device__ __forceinline__ int32_t wide_next(uint32_t h, int4 id, int4 grp, int32_t *stk, int stride, int &top) {
    const uint32_t in = h & ~(static_cast<uint32_t>(grp.z) >> 4);  // internal entries that passed
    const bool i0 = (in & 1u) != 0, i1 = (in & 2u) != 0, i2 = (in & 4u) != 0, i3 = (in & 8u) != 0;
    const int32_t itemL = (i0 && i1) ? grp.x : (i0 ? id.x : (i1 ? id.y : -1));
    const int32_t itemR = (i2 && i3) ? grp.y : (i2 ? id.z : (i3 ? id.w : -1));
    int32_t nx = itemR;
    if (itemL >= 0) {
        if (nx < 0) {
            nx = itemL;
        } else if (top < kStack) {
            stk[top * stride] = itemL;
            ++top;
        }
    }
    if (nx < 0 && top > 0) {
        --top;
        nx = stk[top * stride];
    }
    return nx;
}

// One visit of node `cur`. Returns true when a triangle is hit; otherwise *next gets the node to
// visit next (or -1 when the stack is empty), after pushing the left side's item if both sides
// yield one.
template <int kStack, bool kQuick>
__device__ __forceinline__ bool wide_visit_k(const WSeg &ws, const WNode *__restrict__ wide,
                                           const Node *__restrict__ nodes, const float4 *__restrict__ tris,
                                           int32_t cur, int32_t *next, int32_t *stk, int stride, int &top,
                                           uint32_t &nv, uint32_t &nt) {
    int4 id, grp;
    const uint32_t h = wide_pass<kQuick>(ws, wide, nodes, cur, id, grp);
    ++nv;
    for (uint32_t leaves = wide_leaves(h, grp); leaves; leaves &= leaves - 1u) {
        ++nt;
        if (ray_tri(tris, ~wide_entry(id, static_cast<uint32_t>(__builtin_ctz(leaves))), ws.s.o, ws.s.d)) return true;
    }
    *next = wide_next<kStack>(h, id, grp, stk, stride, top);
    return false;
}

// The visit with the quick box tests when every active lane's ray allows them (a wave-wide
// choice, so the box tests run without per-box branches), else with the checked ones.
template <int kStack>
__device__ __forceinline__ bool wide_visit(const WSeg &ws, const WNode *__restrict__ wide,
                                           const Node *__restrict__ nodes, const float4 *__restrict__ tris,
                                           int32_t cur, int32_t *next, int32_t *stk, int stride, int &top,
                                           uint32_t &nv, uint32_t &nt) {
    if (__ballot(!ws.quick) == 0ull)
        return wide_visit_k<kStack, true>(ws, wide, nodes, tris, cur, next, stk, stride, top, nv, nt);
    return wide_visit_k<kStack, false>(ws, wide, nodes, tris, cur, next, stk, stride, top, nv, nt);
}

// `occluded` (production loop) over the 4-entry nodes: same answer, about half the visits.
template <int kStack>
__device__ __forceinline__ bool occluded_w(const WNode *__restrict__ wide, const Node *__restrict__ nodes,
                                           const float4 *__restrict__ tris, bool treeNested, float maxCoord,
                                           f3 o, f3 d, int32_t *stk, int stride, TraceStats &st) {
    const WSeg s = make_wseg(o, d, treeNested, maxCoord);
    int top = 0;
    int32_t cur = 0;
    uint32_t nv = 0, nt = 0;
    bool hit = false;
    while (cur >= 0) {
        if (wide_visit<kStack>(s, wide, nodes, tris, cur, &cur, stk, stride, top, nv, nt)) {
            hit = true;
            break;
        }
    }
    st.visits += nv;
    st.tests += nt;
    return hit;
}

// Step-wise form for the interleaving schedules: one visit (and the tests of the triangle leaves
// it finds) per step; true when the ray is done, w.hit then has the answer.
struct WWalk {
    WSeg s;
    int32_t cur;  // node to visit next, -1 when the stack is empty
    int top;
    bool hit;
    uint32_t nv, nt;
};
__device__ __forceinline__ void wwalk_begin(WWalk &w, f3 o, f3 d, bool treeNested, float maxCoord) {
    w.s = make_wseg(o, d, treeNested, maxCoord);
    w.cur = 0;
    w.top = 0;
    w.hit = false;
    w.nv = w.nt = 0;
}
template <int kStack>
__device__ __forceinline__ bool walk_step_w(WWalk &w, const WNode *__restrict__ wide, const Node *__restrict__ nodes,
                                            const float4 *__restrict__ tris, int32_t *stk, int stride) {
    int32_t next;
    if (wide_visit<kStack>(w.s, wide, nodes, tris, w.cur, &next, stk, stride, w.top, w.nv, w.nt)) {
        w.hit = true;
        return true;
    }
    w.cur = next;
    return next < 0;
}

// A visit that leaves its triangle tests to the caller, for tracers that gather the tests of a
// wave's rays and run them with the wave's lanes full: the visit of w.cur as walk_step_w makes
// it (same boxes, same pushes, same next node), except that the triangles whose entries pass are
// returned (bit k: entry k, triangle ~ids[k]) instead of tested. A segment is occluded exactly
// when one of the triangles its visits return is hit, so the tests may come later and in any
// order; the walk goes on past a leaf that would have ended it until the caller knows.
template <int kStack>
__device__ __forceinline__ uint32_t walk_step_defer(WWalk &w, const WNode *__restrict__ wide,
                                                   const Node *__restrict__ nodes, int32_t *stk, int stride,
                                                   int4 &ids) {
    int4 grp;
    const uint32_t h = __ballot(!w.s.quick) == 0ull ? wide_pass<true>(w.s, wide, nodes, w.cur, ids, grp)
                                                    : wide_pass<false>(w.s, wide, nodes, w.cur, ids, grp);
    ++w.nv;
    w.cur = wide_next<kStack>(h, ids, grp, stk, stride, w.top);
    return wide_leaves(h, grp);
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
