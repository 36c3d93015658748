// bvh_build.hip — AabbTree::build (src/aabbTreeBuilder.cpp:52-214) on the GPU: the reference's
// 12-bucket binned-SAH tree, node for node and bit for bit, built one tree level per round of
// launches.
//
// The reference pops index ranges of one primitive-order array from a FIFO, so node ids are
// breadth-first and each tree level is a list of disjoint ranges in position order. One round
// handles a whole level:
//   classify + scan    ranges of >= 2 primitives take the next node ids (breadth-first order),
//                      ranges of >= 3 are split; 1- and 2-primitive ranges become leaves
//   bounds             centroid and box bounds of each split range
//   bin                bucket of each primitive on the range's longest centroid axis; box and
//                      count per bucket
//   choose             the reference's SAH sweep over the 12 buckets, one lane per range
//   partition          the reference's swap partition, reproduced exactly (below)
// and the host reads back only the level's node and split counts.
//
// Exact min/max. nv_min(a, b) = a < b ? a : b (nvmath.inl:2414-2420) folded in primitive order
// returns, among equal values, the LAST one; equal floats differ in bits only as +0/-0. Each fold
// is therefore an order-independent atomic min (max) over 64-bit keys: the value in a monotone
// integer code with -0 folded onto +0, then the position (complemented for min) so that the
// later position wins a tie; the winner's own bits are then read back at that position.
//
// Exact partition. The reference swaps every left primitive (bucket <= split) with the first
// element of the right block (aabbTreeBuilder.cpp:178-183). Left primitives keep their order;
// the right block acts as a queue whose front moves to its back whenever a left primitive
// arrives. After the pass, a position p >= mid holds what was placed there at step p: the
// primitive itself if it is a right one, else the front of the block at that step, which sat at
// first + (left primitives before p) and is resolved the same way. That chain only moves to
// smaller positions; it is followed by pointer jumping, ceil(log2(range length)) rounds.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstdint>
#include <cstring>

#include "ctx.h"
#include "kernels.h"
#include "scan.h"

using namespace rki;

namespace {

constexpr int kBuckets = 12;
constexpr int kBoundWords = 12;                 // centroid lo/hi, box lo/hi keys
constexpr int kBinWords = 7;                    // box lo/hi keys + count
constexpr int kAcc = kBoundWords + kBuckets * kBinWords;  // 64-bit words per split range
constexpr int kSlots = 4;  // split ranges a 256-lane group pre-reduces in LDS

struct Seg {  // a range of the level; link = parent node * 2 + side, -1 for the root
    uint32_t first, last;
    int32_t link, pad;
};
struct Split {  // a range of >= 3 primitives being split at this level
    uint32_t id, first, last, mid;
    uint32_t bucket;      // last bucket of the left side
    uint32_t degenerate;  // the sweep put everything on one side: halved instead
    uint32_t pad[2];
};
struct Params {
    int32_t axis;
    float lo, width, parentArea;
};
struct Info {
    unsigned long long totals;  // node count (low 32 bits) and split count (high 32 bits)
    uint32_t maxLen;            // longest split range of the level
    uint32_t bad;               // a vertex coordinate is not finite
};

using u64 = unsigned long long;

__device__ inline float nv_min(float a, float b) { return a < b ? a : b; }
__device__ inline float nv_max(float a, float b) { return a > b ? a : b; }

__device__ inline uint32_t code(float v) {
    const uint32_t u = __float_as_uint(v == 0.0f ? 0.0f : v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float uncode(uint32_t c) {
    return __uint_as_float((c & 0x80000000u) ? (c & 0x7fffffffu) : ~c);
}
__device__ inline u64 key_min(float v, uint32_t p) { return (static_cast<u64>(code(v)) << 32) | static_cast<uint32_t>(~p); }
__device__ inline u64 key_max(float v, uint32_t p) { return (static_cast<u64>(code(v)) << 32) | p; }

// word layout of a split range's accumulator: [0,3) centroid lo, [3,6) centroid hi, [6,9) box
// lo, [9,12) box hi; then per bucket b at 12 + 7b: [0,3) box lo, [3,6) box hi, [6] count
__host__ __device__ inline bool is_min_word(int w) { return w < kBoundWords ? ((w / 3) & 1) == 0 : ((w - kBoundWords) % kBinWords) < 3; }
__host__ __device__ inline bool is_count_word(int w) { return w >= kBoundWords && (w - kBoundWords) % kBinWords == 6; }
__device__ inline u64 word_init(int w) { return (!is_count_word(w) && is_min_word(w)) ? ~0ull : 0ull; }

__device__ inline void atomic_key(u64 *a, u64 k, bool isMin) {
    if (isMin) atomicMin(a, k);
    else atomicMax(a, k);
}

// component k of position p's box lo (which 1), box hi (2) or centroid (0)
__device__ inline float element(const uint32_t *idx, const float4 *box, uint32_t p, int which, int k) {
    const uint32_t t = idx[p];
    const float4 lo = box[2 * t], hi = box[2 * t + 1];
    const float l = k == 0 ? lo.x : (k == 1 ? lo.y : lo.z);
    const float h = k == 0 ? hi.x : (k == 1 ? hi.y : hi.z);
    return which == 1 ? l : (which == 2 ? h : 0.5f * (l + h));
}
__device__ inline float decode(u64 key, bool isMin, const uint32_t *idx, const float4 *box, int which, int k) {
    const float v = uncode(static_cast<uint32_t>(key >> 32));
    if (v != 0.0f) return v;
    const uint32_t p = isMin ? ~static_cast<uint32_t>(key) : static_cast<uint32_t>(key);
    return element(idx, box, p, which, k);  // +0 or -0: the winner's own bits
}

__device__ inline float half_area(const float lo[3], const float hi[3]) {  // surfaceAreaHeuristic
    const float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
    return sx * sy + sx * sz + sy * sz;
}

struct Bin {  // Bucket (aabbTreeBuilder.cpp:33-50)
    float lo[3], hi[3];
    u64 count;
};
__device__ inline void bin_empty(Bin &b) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = FLT_MAX;
        b.hi[k] = -FLT_MAX;
    }
    b.count = 0;
}
__device__ inline Bin bin_merged(Bin a, const Bin &b) {
    a.count += b.count;
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = nv_min(a.lo[k], b.lo[k]);
        a.hi[k] = nv_max(a.hi[k], b.hi[k]);
    }
    return a;
}
__device__ inline float bin_cost(const Bin &b) { return static_cast<float>(b.count) * half_area(b.lo, b.hi); }

__device__ inline void write_box(float *dst, const float *v) {  // vec4(vec3) sets w = 1
    dst[0] = v[0];
    dst[1] = v[1];
    dst[2] = v[2];
    dst[3] = 1.0f;
}

// ---- kernels ----

// triangles out (w = 1), primitive boxes (aabbForTriangle), the identity order, one root range
__global__ __launch_bounds__(256) void k_bvh_prep(uint32_t n, const float *xyz, restir_triangle *tris,
                                                  float4 *box, uint32_t *idx, int32_t *segOf, Info *info) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float v[9];
    bool finite = true;
    for (int k = 0; k < 9; ++k) {
        v[k] = xyz[9 * static_cast<size_t>(i) + k];
        finite = finite && (v[k] - v[k] == 0.0f);
    }
    if (!finite) atomicOr(&info->bad, 1u);
    restir_triangle t;
    for (int k = 0; k < 3; ++k) {
        t.p1[k] = v[k];
        t.p2[k] = v[3 + k];
        t.p3[k] = v[6 + k];
    }
    t.p1[3] = t.p2[3] = t.p3[3] = 1.0f;
    tris[i] = t;
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = v[k];
        hi[k] = v[k];
        lo[k] = nv_min(lo[k], v[3 + k]);
        hi[k] = nv_max(hi[k], v[3 + k]);
        lo[k] = nv_min(lo[k], v[6 + k]);
        hi[k] = nv_max(hi[k], v[6 + k]);
    }
    box[2 * i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    box[2 * i + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    idx[i] = i;
    segOf[i] = 0;
}

__global__ __launch_bounds__(256) void k_bvh_classify(uint32_t S, const Seg *segs, u64 *flags, Info *info) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= S) return;
    const uint32_t len = segs[s].last - segs[s].first;
    flags[s] = static_cast<u64>(len >= 2) | (static_cast<u64>(len >= 3) << 32);
    if (len >= 3) atomicMax(&info->maxLen, len);
}

// node ids, links from the parents, leaf pairs, the split list
__global__ __launch_bounds__(256) void k_bvh_setup(uint32_t S, const Seg *segs, const u64 *scan, uint32_t nodeBase,
                                                   const uint32_t *idx, const float4 *box, restir_aabb_tree_node *nodes,
                                                   Split *splits) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= S) return;
    const Seg g = segs[s];
    const uint32_t len = g.last - g.first;
    const uint32_t id = nodeBase + static_cast<uint32_t>(scan[s]);
    const int32_t value = len == 1 ? ~static_cast<int32_t>(idx[g.first]) : static_cast<int32_t>(id);
    if (g.link >= 0) {
        restir_aabb_tree_node &parent = nodes[g.link >> 1];
        if (g.link & 1) parent.rightChild = value;
        else parent.leftChild = value;
    }
    if (len == 2) {
        const uint32_t a = idx[g.first], b = idx[g.first + 1];
        restir_aabb_tree_node node;
        const float4 alo = box[2 * a], ahi = box[2 * a + 1], blo = box[2 * b], bhi = box[2 * b + 1];
        const float la[3] = {alo.x, alo.y, alo.z}, ha[3] = {ahi.x, ahi.y, ahi.z};
        const float lb[3] = {blo.x, blo.y, blo.z}, hb[3] = {bhi.x, bhi.y, bhi.z};
        write_box(node.leftAabbMin, la);
        write_box(node.leftAabbMax, ha);
        write_box(node.rightAabbMin, lb);
        write_box(node.rightAabbMax, hb);
        node.leftChild = ~static_cast<int32_t>(a);
        node.rightChild = ~static_cast<int32_t>(b);
        node._pad[0] = node._pad[1] = 0;
        nodes[id] = node;
    } else if (len >= 3) {
        Split sp{};
        sp.id = id;
        sp.first = g.first;
        sp.last = g.last;
        splits[scan[s] >> 32] = sp;
    }
}

__global__ __launch_bounds__(256) void k_bvh_init_acc(uint32_t words, u64 *acc) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < words) acc[i] = word_init(static_cast<int>(i % kAcc));
}

// rank of the split range holding each position (-1: a finished or 1-2 primitive range)
__global__ __launch_bounds__(256) void k_bvh_rank(uint32_t n, const int32_t *segOf, const Seg *segs, const u64 *scan,
                                                  int32_t *rank) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= n) return;
    const int32_t s = segOf[p];
    int32_t r = -1;
    if (s >= 0 && segs[s].last - segs[s].first >= 3) r = static_cast<int32_t>(scan[s] >> 32);
    rank[p] = r;
}

// Adds `nw` keys (words w0.. of range r's accumulator) through LDS when r is one of the first
// kSlots ranges present in the group (ranges are position-ordered, so a group's ranks are
// consecutive), else straight to memory. Counts add; other words take min or max.
template <int NW>
struct GroupAcc {
    u64 *sh;  // kSlots * NW words
    int r0;
    __device__ void add(u64 *acc, int r, int w0, const u64 *keys, int nKeys) {
        const int slot = r - r0;
        for (int j = 0; j < nKeys; ++j) {
            const int w = w0 + j;
            if (slot < kSlots) {
                u64 *a = &sh[slot * NW + (w - base)];
                if (is_count_word(w)) atomicAdd(a, keys[j]);
                else atomic_key(a, keys[j], is_min_word(w));
            } else {
                u64 *a = &acc[static_cast<size_t>(r) * kAcc + w];
                if (is_count_word(w)) atomicAdd(a, keys[j]);
                else atomic_key(a, keys[j], is_min_word(w));
            }
        }
    }
    int base;  // first accumulator word the LDS copy mirrors
};

__device__ inline int group_min_rank(int r, int *shR) {
    if (threadIdx.x == 0) *shR = INT_MAX;
    __syncthreads();
    if (r >= 0) atomicMin(shR, r);
    __syncthreads();
    return *shR;
}

// flushes the LDS copies (words base.. base+NW of ranges r0.. r0+kSlots) to memory
template <int NW>
__device__ inline void group_flush(const u64 *sh, int r0, int base, u64 *acc) {
    __syncthreads();
    if (r0 == INT_MAX) return;
    for (int i = threadIdx.x; i < kSlots * NW; i += 256) {
        const int w = base + i % NW;
        const u64 v = sh[i];
        if (v == word_init(w)) continue;  // nothing added (empty min/max word or zero count)
        u64 *a = &acc[static_cast<size_t>(r0 + i / NW) * kAcc + w];
        if (is_count_word(w)) atomicAdd(a, v);
        else atomic_key(a, v, is_min_word(w));
    }
}

__global__ __launch_bounds__(256) void k_bvh_bounds(uint32_t n, const uint32_t *idx, const int32_t *rank, const float4 *box,
                                                    u64 *acc) {
    __shared__ u64 sh[kSlots * kBoundWords];
    __shared__ int shR;
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    const int r = p < n ? rank[p] : -1;
    if (threadIdx.x < kSlots * kBoundWords) sh[threadIdx.x] = word_init(threadIdx.x % kBoundWords);
    const int r0 = group_min_rank(r, &shR);
    if (r >= 0) {
        const uint32_t t = idx[p];
        const float4 lo = box[2 * t], hi = box[2 * t + 1];
        const float cx = 0.5f * (lo.x + hi.x), cy = 0.5f * (lo.y + hi.y), cz = 0.5f * (lo.z + hi.z);
        const u64 keys[kBoundWords] = {key_min(cx, p),   key_min(cy, p),   key_min(cz, p),   key_max(cx, p),
                                       key_max(cy, p),   key_max(cz, p),   key_min(lo.x, p), key_min(lo.y, p),
                                       key_min(lo.z, p), key_max(hi.x, p), key_max(hi.y, p), key_max(hi.z, p)};
        GroupAcc<kBoundWords> g{sh, r0, 0};
        g.add(acc, r, 0, keys, kBoundWords);
    }
    group_flush<kBoundWords>(sh, r0, 0, acc);
}

__global__ __launch_bounds__(256) void k_bvh_params(uint32_t nSplits, const uint32_t *idx, const float4 *box,
                                                    const u64 *acc, Params *par) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= nSplits) return;
    const u64 *a = acc + static_cast<size_t>(r) * kAcc;
    float cLo[3], cHi[3], bLo[3], bHi[3];
    for (int k = 0; k < 3; ++k) {
        cLo[k] = decode(a[k], true, idx, box, 0, k);
        cHi[k] = decode(a[3 + k], false, idx, box, 0, k);
        bLo[k] = decode(a[6 + k], true, idx, box, 1, k);
        bHi[k] = decode(a[9 + k], false, idx, box, 2, k);
    }
    Params q;
    q.parentArea = half_area(bLo, bHi);
    float extent[3];
    for (int k = 0; k < 3; ++k) extent[k] = cHi[k] - cLo[k];
    int axis = extent[0] > extent[1] ? 0 : 1;
    if (extent[2] > extent[axis]) axis = 2;
    q.axis = axis;
    q.lo = cLo[axis];
    q.width = extent[axis] / static_cast<float>(kBuckets);
    par[r] = q;
}

__global__ __launch_bounds__(256) void k_bvh_bin(uint32_t n, const uint32_t *idx, const int32_t *rank, const float4 *box,
                                                 const Params *par, uint8_t *bucket, u64 *acc) {
    constexpr int NW = kBuckets * kBinWords;
    __shared__ u64 sh[kSlots * NW];
    __shared__ int shR;
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    const int r = p < n ? rank[p] : -1;
    for (int i = threadIdx.x; i < kSlots * NW; i += 256) sh[i] = word_init(kBoundWords + i % NW);
    const int r0 = group_min_rank(r, &shR);
    if (r >= 0) {
        const Params q = par[r];
        const uint32_t t = idx[p];
        const float4 lo = box[2 * t], hi = box[2 * t + 1];
        const float c = q.axis == 0 ? 0.5f * (lo.x + hi.x) : (q.axis == 1 ? 0.5f * (lo.y + hi.y) : 0.5f * (lo.z + hi.z));
        float f = (c - q.lo) / q.width;
        f = f < 0.5f ? 0.5f : f;  // nv_clamp(f, 0.5, numBuckets - 0.5)
        f = f > static_cast<float>(kBuckets) - 0.5f ? static_cast<float>(kBuckets) - 0.5f : f;
        // coincident centroids give 0/0 (undefined in the reference): bucket 0
        const uint32_t b = f == f ? static_cast<uint32_t>(f) : 0u;
        bucket[p] = static_cast<uint8_t>(b);
        const u64 keys[kBinWords] = {key_min(lo.x, p), key_min(lo.y, p), key_min(lo.z, p), key_max(hi.x, p),
                                     key_max(hi.y, p), key_max(hi.z, p), 1ull};
        GroupAcc<NW> g{sh, r0, kBoundWords};
        g.add(acc, r, kBoundWords + static_cast<int>(b) * kBinWords, keys, kBinWords);
    }
    group_flush<NW>(sh, r0, kBoundWords, acc);
}

// the SAH sweep of one range (aabbTreeBuilder.cpp:131-175), the split point, the next level's
// two ranges
__global__ __launch_bounds__(256) void k_bvh_choose(uint32_t nSplits, const uint32_t *idx, const float4 *box,
                                                    const Params *par, u64 *acc, Split *splits,
                                                    restir_aabb_tree_node *nodes, Seg *next) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= nSplits) return;
    u64 *a = acc + static_cast<size_t>(r) * kAcc;
    Bin bins[kBuckets];
#pragma unroll
    for (int b = 0; b < kBuckets; ++b) {
        const u64 *w = a + kBoundWords + b * kBinWords;
        bin_empty(bins[b]);
        bins[b].count = w[6];
        if (w[6] == 0) continue;
        for (int k = 0; k < 3; ++k) {  // the bucket's fold starts from the empty box
            bins[b].lo[k] = nv_min(FLT_MAX, decode(w[k], true, idx, box, 1, k));
            bins[b].hi[k] = nv_max(-FLT_MAX, decode(w[3 + k], false, idx, box, 2, k));
        }
    }
    Bin right[kBuckets - 1];
    {
        Bin cur = bins[kBuckets - 1];
#pragma unroll
        for (int s = kBuckets - 1; s > 0;) {
            right[--s] = cur;
            cur = bin_merged(cur, bins[s]);
        }
    }
    const float parentArea = par[r].parentArea;
    uint32_t split = 0;
    float lLo[3] = {0, 0, 0}, lHi[3] = {0, 0, 0}, rLo[3] = {0, 0, 0}, rHi[3] = {0, 0, 0};
    {
        float best = FLT_MAX;
        Bin left;
        bin_empty(left);
#pragma unroll
        for (int s = 0; s < kBuckets - 1; ++s) {
            left = bin_merged(left, bins[s]);
            const float cost = 0.125f + (bin_cost(left) + bin_cost(right[s])) / parentArea;
            if (cost < best) {
                best = cost;
                split = static_cast<uint32_t>(s);
                for (int k = 0; k < 3; ++k) {
                    lLo[k] = left.lo[k];
                    lHi[k] = left.hi[k];
                    rLo[k] = right[s].lo[k];
                    rHi[k] = right[s].hi[k];
                }
            }
        }
    }
    Split sp = splits[r];
    u64 nLeft = 0;
    for (int b = 0; b < kBuckets; ++b)
        if (static_cast<uint32_t>(b) <= split) nLeft += bins[b].count;
    uint32_t mid = sp.first + static_cast<uint32_t>(nLeft);
    const bool degenerate = mid == sp.first || mid == sp.last;
    if (degenerate) {
        mid = static_cast<uint32_t>((static_cast<uint64_t>(sp.first) + sp.last) / 2);
        for (int w = 0; w < kBoundWords; ++w) a[w] = word_init(w);  // reused for the two halves' boxes
    }
    restir_aabb_tree_node &node = nodes[sp.id];
    if (!degenerate) {
        write_box(node.leftAabbMin, lLo);
        write_box(node.leftAabbMax, lHi);
        write_box(node.rightAabbMin, rLo);
        write_box(node.rightAabbMax, rHi);
    }
    node._pad[0] = node._pad[1] = 0;
    sp.mid = mid;
    sp.bucket = split;
    sp.degenerate = degenerate ? 1u : 0u;
    splits[r] = sp;
    next[2 * r] = Seg{sp.first, mid, static_cast<int32_t>(2 * sp.id), 0};
    next[2 * r + 1] = Seg{mid, sp.last, static_cast<int32_t>(2 * sp.id + 1), 0};
}

// boxes of the two halves of a range the sweep could not split (aabbTreeBuilder.cpp:186-203)
__global__ __launch_bounds__(256) void k_bvh_halves(uint32_t n, const uint32_t *idx, const int32_t *rank, const float4 *box,
                                                    const Split *splits, u64 *acc) {
    __shared__ u64 sh[kSlots * kBoundWords];
    __shared__ int shR;
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    int r = p < n ? rank[p] : -1;
    if (r >= 0 && !splits[r].degenerate) r = -1;
    if (threadIdx.x < kSlots * kBoundWords) sh[threadIdx.x] = word_init(threadIdx.x % kBoundWords);
    const int r0 = group_min_rank(r, &shR);
    if (r >= 0) {
        const uint32_t t = idx[p];
        const float4 lo = box[2 * t], hi = box[2 * t + 1];
        const u64 keys[6] = {key_min(lo.x, p), key_min(lo.y, p), key_min(lo.z, p),
                             key_max(hi.x, p), key_max(hi.y, p), key_max(hi.z, p)};
        GroupAcc<kBoundWords> g{sh, r0, 0};
        g.add(acc, r, p < splits[r].mid ? 0 : 6, keys, 6);
    }
    group_flush<kBoundWords>(sh, r0, 0, acc);
}

__global__ __launch_bounds__(256) void k_bvh_halves_write(uint32_t nSplits, const uint32_t *idx, const float4 *box,
                                                          const Split *splits, const u64 *acc,
                                                          restir_aabb_tree_node *nodes) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= nSplits || !splits[r].degenerate) return;
    const u64 *a = acc + static_cast<size_t>(r) * kAcc;
    float v[4][3];
    for (int h = 0; h < 4; ++h)  // left lo, left hi, right lo, right hi
        for (int k = 0; k < 3; ++k) v[h][k] = decode(a[3 * h + k], (h & 1) == 0, idx, box, (h & 1) ? 2 : 1, k);
    restir_aabb_tree_node &node = nodes[splits[r].id];
    write_box(node.leftAabbMin, v[0]);
    write_box(node.leftAabbMax, v[1]);
    write_box(node.rightAabbMin, v[2]);
    write_box(node.rightAabbMax, v[3]);
}

__device__ inline bool goes_left(const Split &sp, const uint8_t *bucket, uint32_t p) {
    return !sp.degenerate && bucket[p] <= sp.bucket;
}

__global__ __launch_bounds__(256) void k_bvh_left_flags(uint32_t n, const int32_t *rank, const Split *splits,
                                                        const uint8_t *bucket, uint32_t *flag) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= n) return;
    const int r = rank[p];
    flag[p] = (r >= 0 && goes_left(splits[r], bucket, p)) ? 1u : 0u;
}

// ptr[p]: the position whose placement position p copies (itself when final)
__global__ __launch_bounds__(256) void k_bvh_chain(uint32_t n, const int32_t *rank, const Split *splits,
                                                   const uint8_t *bucket, const uint32_t *leftBefore, uint32_t *ptr) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= n) return;
    const int r = rank[p];
    uint32_t q = p;
    if (r >= 0) {
        const Split sp = splits[r];
        if (goes_left(sp, bucket, p)) {
            const uint32_t g = sp.first + (leftBefore[p] - leftBefore[sp.first]);
            q = g < p ? g : p;
        }
    }
    ptr[p] = q;
}

__global__ __launch_bounds__(256) void k_bvh_jump(uint32_t n, const uint32_t *in, uint32_t *out) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p < n) out[p] = in[in[p]];
}

// the level's new primitive order and each position's range at the next level
__global__ __launch_bounds__(256) void k_bvh_scatter(uint32_t n, const int32_t *rank, const Split *splits,
                                                     const uint8_t *bucket, const uint32_t *leftBefore,
                                                     const uint32_t *src, const uint32_t *idx, uint32_t *idxOut,
                                                     int32_t *segOut) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= n) return;
    const int r = rank[p];
    if (r < 0) {
        idxOut[p] = idx[p];
        segOut[p] = -1;
        return;
    }
    const Split sp = splits[r];
    if (sp.degenerate) {
        idxOut[p] = idx[p];
        segOut[p] = 2 * r + (p >= sp.mid ? 1 : 0);
        return;
    }
    if (goes_left(sp, bucket, p)) {
        const uint32_t q = sp.first + (leftBefore[p] - leftBefore[sp.first]);
        idxOut[q] = idx[p];
        segOut[q] = 2 * r;
    }
    if (p >= sp.mid) {
        idxOut[p] = idx[src[p]];
        segOut[p] = 2 * r + 1;
    }
}

inline uint32_t groups(uint64_t n) { return static_cast<uint32_t>((n + 255) / 256); }

inline size_t align_up(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

struct Layout {
    size_t box, idx0, idx1, seg0, seg1, rank, bucket, flag, ptr0, ptr1, temp32, segs0, segs1, flags64, temp64, splits,
        params, acc, info, total;
};

Layout layout(uint32_t n) {
    const size_t maxSplits = n / 3 + 1, maxSegs = 2 * maxSplits + 1;
    Layout L{};
    size_t o = 0;
    auto take = [&o](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes);
        return at;
    };
    L.box = take(sizeof(float4) * 2 * n);
    L.idx0 = take(4ull * n);
    L.idx1 = take(4ull * n);
    L.seg0 = take(4ull * n);
    L.seg1 = take(4ull * n);
    L.rank = take(4ull * n);
    L.bucket = take(n);
    L.flag = take(4ull * n);
    L.ptr0 = take(4ull * n);
    L.ptr1 = take(4ull * n);
    L.temp32 = take(4 * rscan::temp_items(n));
    L.segs0 = take(sizeof(Seg) * maxSegs);
    L.segs1 = take(sizeof(Seg) * maxSegs);
    L.flags64 = take(8 * maxSegs);
    L.temp64 = take(8 * rscan::temp_items(static_cast<uint32_t>(maxSegs)));
    L.splits = take(sizeof(Split) * maxSplits);
    L.params = take(sizeof(Params) * maxSplits);
    L.acc = take(8ull * kAcc * maxSplits);
    L.info = take(sizeof(Info));
    L.total = o;
    return L;
}

}  // namespace

// ------------------------------------------------------------ 4-entry nodes of a bound tree --
namespace {

__device__ inline bool finite_box(float4 mn, float4 mx) {
    constexpr float kInf = __builtin_huge_valf();
    return fabsf(mn.x) < kInf && fabsf(mn.y) < kInf && fabsf(mn.z) < kInf && fabsf(mx.x) < kInf &&
           fabsf(mx.y) < kInf && fabsf(mx.z) < kInf && mn.x <= mx.x && mn.y <= mx.y && mn.z <= mx.z;
}
__device__ inline float box_max(float4 v) { return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fabsf(v.z)); }
__device__ inline bool inside(float4 mn, float4 mx, float4 outerMn, float4 outerMx) {
    return outerMn.x <= mn.x && outerMn.y <= mn.y && outerMn.z <= mn.z && mx.x <= outerMx.x &&
           mx.y <= outerMx.y && mx.z <= outerMx.z;
}

// One thread per reference node x: its WNode (see rt_traverse.h), and the check that both boxes
// stored in x are finite and ordered and that each internal child's two boxes lie inside the box
// x stores for it (info[0] != 0 otherwise). Every box is stored in exactly one node, so the
// threads together check every box and every parent-child pair; info[1] gets the bits of the
// largest |coordinate| of any box (non-negative floats order as their bit patterns; a NaN orders
// above every number).
__global__ __launch_bounds__(256) void k_wide_build(const rt::Node *__restrict__ nodes, uint32_t nNodes,
                                                    rt::WNode *__restrict__ wide, uint32_t *__restrict__ info) {
    const uint32_t x = blockIdx.x * 256u + threadIdx.x;
    if (x >= nNodes) return;
    const rt::Node b = rt::load_node(nodes, static_cast<int32_t>(x));
    bool ok = finite_box(b.lmin, b.lmax) && finite_box(b.rmin, b.rmax);
    float4 mn[4], mx[4];
    int32_t id[4], grp[2];
    for (int side = 0; side < 2; ++side) {
        const int32_t c = side ? b.right : b.left;
        const float4 smn = side ? b.rmin : b.lmin, smx = side ? b.rmax : b.lmax;
        if (c < 0) {
            mn[2 * side] = smn;
            mx[2 * side] = smx;
            id[2 * side] = c;
            mn[2 * side + 1] = mx[2 * side + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            id[2 * side + 1] = rt::kNoEntry;
            grp[side] = -1;
        } else {
            const rt::Node cn = rt::load_node(nodes, c);
            mn[2 * side] = cn.lmin;
            mx[2 * side] = cn.lmax;
            id[2 * side] = cn.left;
            mn[2 * side + 1] = cn.rmin;
            mx[2 * side + 1] = cn.rmax;
            id[2 * side + 1] = cn.right;
            grp[side] = c;
            ok = ok && inside(cn.lmin, cn.lmax, smn, smx) && inside(cn.rmin, cn.rmax, smn, smx);
        }
    }
    float4 *o = reinterpret_cast<float4 *>(wide + x);
    o[0] = make_float4(mn[0].x, mn[1].x, mn[2].x, mn[3].x);
    o[1] = make_float4(mn[0].y, mn[1].y, mn[2].y, mn[3].y);
    o[2] = make_float4(mn[0].z, mn[1].z, mn[2].z, mn[3].z);
    o[3] = make_float4(mx[0].x, mx[1].x, mx[2].x, mx[3].x);
    o[4] = make_float4(mx[0].y, mx[1].y, mx[2].y, mx[3].y);
    o[5] = make_float4(mx[0].z, mx[1].z, mx[2].z, mx[3].z);
    reinterpret_cast<int4 *>(o)[6] = make_int4(id[0], id[1], id[2], id[3]);
    // entry masks: bit k set when entry k exists, bit 4 + k when it is a triangle leaf
    uint32_t masks = 0;
    for (int k = 0; k < 4; ++k) {
        if (id[k] != rt::kNoEntry) masks |= 1u << k;
        if (id[k] != rt::kNoEntry && id[k] < 0) masks |= 16u << k;
    }
    reinterpret_cast<int4 *>(o)[7] = make_int4(grp[0], grp[1], static_cast<int32_t>(masks), 0);
    if (!ok) atomicOr(&info[0], 1u);
    const float m = fmaxf(fmaxf(fmaxf(box_max(b.lmin), box_max(b.lmax)), box_max(b.rmin)), box_max(b.rmax));
    const bool nan = b.lmin.x != b.lmin.x || b.lmin.y != b.lmin.y || b.lmin.z != b.lmin.z || b.lmax.x != b.lmax.x ||
                     b.lmax.y != b.lmax.y || b.lmax.z != b.lmax.z || b.rmin.x != b.rmin.x || b.rmin.y != b.rmin.y ||
                     b.rmin.z != b.rmin.z || b.rmax.x != b.rmax.x || b.rmax.y != b.rmax.y || b.rmax.z != b.rmax.z;
    atomicMax(&info[1], nan ? 0x7fc00000u : __float_as_uint(m));
}

}  // namespace

namespace rki {
// Binds a validated tree (nNodes = nTris - 1, every index in range, depth <= 64) to the context
// and builds its 4-entry nodes. Waits for the device first: work still in flight may read the
// previous tree's 4-entry nodes, which share the context buffer.
int bind_tree(restir_ctx *ctx, const void *nodes, uint32_t nNodes, const void *tris, uint32_t depth, hipStream_t s) {
    auto hip = [&](hipError_t e, const char *what) { return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, what); };
    if (int r = hip(hipDeviceSynchronize(), "bind tree: wait for work in flight")) return r;
    const size_t wideBytes = sizeof(rt::WNode) * static_cast<size_t>(nNodes);
    if (int r = ensure(ctx, &ctx->wide, &ctx->wideBytes, wideBytes + 256)) return r;
    rt::WNode *wide = static_cast<rt::WNode *>(ctx->wide);
    uint32_t *info = reinterpret_cast<uint32_t *>(static_cast<char *>(ctx->wide) + wideBytes);
    if (int r = hip(hipMemsetAsync(info, 0, 2 * sizeof(uint32_t), s), "bind tree: clear")) return r;
    RK_LAUNCH("k_wide_build", k_wide_build, dim3(groups(nNodes)), dim3(256), 0, s,
              static_cast<const rt::Node *>(nodes), nNodes, wide, info);
    uint32_t host[2] = {1u, 0x7fc00000u};
    if (int r = hip(hipMemcpyAsync(host, info, sizeof host, hipMemcpyDeviceToHost, s), "bind tree: read")) return r;
    if (int r = hip(hipStreamSynchronize(s), "bind tree: finish")) return r;
    float maxCoord;
    std::memcpy(&maxCoord, &host[1], sizeof maxCoord);
    ctx->scene.nodes = static_cast<const rt::Node *>(nodes);
    ctx->scene.tris = static_cast<const float4 *>(tris);
    ctx->scene.wide = wide;
    ctx->scene.treeNested = host[0] == 0 ? 1 : 0;
    ctx->scene.maxCoord = maxCoord;
    // a walk defers at most one sibling per internal level below the root: depth - 1 entries
    ctx->stack = depth <= 24 ? 24 : (depth <= 32 ? 32 : 64);
    ctx->persistentBlocks = persistent_blocks(ctx->cus, ctx->stack);
    ctx->treeDepth = depth;
    ctx->bvhSet = true;
    return RESTIR_OK;
}
}  // namespace rki

extern "C" int restir_build_bvh(restir_ctx *ctx, const float *triXyz, uint32_t nTris, restir_aabb_tree_node *nodes,
                                restir_triangle *tris, uint32_t *depthOut, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    ctx->bvhSet = false;
    if (!triXyz || !nodes || !tris) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "build_bvh: NULL buffer");
    if (nTris < 2) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "build_bvh: %u triangles (need >= 2)", nTris);
    if (nTris > (1u << 30)) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "build_bvh: %u triangles > 2^30", nTris);
    const size_t xyzBytes = 36ull * nTris, nodeBytes = sizeof(restir_aabb_tree_node) * (nTris - 1ull),
                 triBytes = sizeof(restir_triangle) * static_cast<size_t>(nTris);
    if (overlaps(triXyz, xyzBytes, nodes, nodeBytes) || overlaps(triXyz, xyzBytes, tris, triBytes) ||
        overlaps(nodes, nodeBytes, tris, triBytes))
        return fail(ctx, RESTIR_E_ALIASING, "build_bvh: vertex, node and triangle buffers must not overlap");
    if (int r = use_device(ctx)) return r;
    const Layout L = layout(nTris);
    if (int r = ensure(ctx, &ctx->bvhScratch, &ctx->bvhScratchBytes, L.total)) return r;
    if (!ctx->bvhInfoHost) {
        const hipError_t e = hipHostMalloc(&ctx->bvhInfoHost, sizeof(Info), hipHostMallocDefault);
        if (e != hipSuccess) {
            ctx->bvhInfoHost = nullptr;
            return hip_fail(ctx, e, "build_bvh: hipHostMalloc");
        }
    }
    char *base = static_cast<char *>(ctx->bvhScratch);
    float4 *box = reinterpret_cast<float4 *>(base + L.box);
    uint32_t *idx[2] = {reinterpret_cast<uint32_t *>(base + L.idx0), reinterpret_cast<uint32_t *>(base + L.idx1)};
    int32_t *segOf[2] = {reinterpret_cast<int32_t *>(base + L.seg0), reinterpret_cast<int32_t *>(base + L.seg1)};
    int32_t *rank = reinterpret_cast<int32_t *>(base + L.rank);
    uint8_t *bucket = reinterpret_cast<uint8_t *>(base + L.bucket);
    uint32_t *flag = reinterpret_cast<uint32_t *>(base + L.flag);
    uint32_t *ptr[2] = {reinterpret_cast<uint32_t *>(base + L.ptr0), reinterpret_cast<uint32_t *>(base + L.ptr1)};
    uint32_t *temp32 = reinterpret_cast<uint32_t *>(base + L.temp32);
    Seg *segs[2] = {reinterpret_cast<Seg *>(base + L.segs0), reinterpret_cast<Seg *>(base + L.segs1)};
    u64 *flags64 = reinterpret_cast<u64 *>(base + L.flags64);
    u64 *temp64 = reinterpret_cast<u64 *>(base + L.temp64);
    Split *splits = reinterpret_cast<Split *>(base + L.splits);
    Params *params = reinterpret_cast<Params *>(base + L.params);
    u64 *acc = reinterpret_cast<u64 *>(base + L.acc);
    Info *info = reinterpret_cast<Info *>(base + L.info);
    Info *host = static_cast<Info *>(ctx->bvhInfoHost);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t n = nTris;

    auto hip = [&](hipError_t e, const char *what) { return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, what); };
    if (int r = hip(hipMemsetAsync(info, 0, sizeof(Info), s), "build_bvh: clear")) return r;
    RK_LAUNCH("k_bvh_prep", k_bvh_prep, dim3(groups(n)), dim3(256), 0, s, n, triXyz, tris, box, idx[0], segOf[0], info);
    const Seg root{0u, n, -1, 0};
    if (int r = hip(hipMemcpyAsync(segs[0], &root, sizeof(Seg), hipMemcpyHostToDevice, s), "build_bvh: root")) return r;

    uint32_t S = 1, nodeBase = 0, depth = 0, cur = 0;
    for (uint32_t level = 1; S > 0; ++level) {
        if (int r = hip(hipMemsetAsync(&info->maxLen, 0, sizeof(uint32_t), s), "build_bvh: clear")) return r;
        RK_LAUNCH("k_bvh_classify", k_bvh_classify, dim3(groups(S)), dim3(256), 0, s, S, segs[cur], flags64, info);
        if (int r = hip(rscan::exclusive_scan<u64>(flags64, flags64, S, temp64, &info->totals, s), "build_bvh: scan"))
            return r;
        if (int r = hip(hipMemcpyAsync(host, info, sizeof(Info), hipMemcpyDeviceToHost, s), "build_bvh: read counts"))
            return r;
        if (int r = hip(hipStreamSynchronize(s), "build_bvh: level")) return r;
        if (host->bad) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "build_bvh: a vertex coordinate is not finite");
        const uint32_t nodesHere = static_cast<uint32_t>(host->totals & 0xffffffffu);
        const uint32_t nSplits = static_cast<uint32_t>(host->totals >> 32);
        if (nodesHere > 0) depth = level;
        RK_LAUNCH("k_bvh_setup", k_bvh_setup, dim3(groups(S)), dim3(256), 0, s, S, segs[cur], flags64, nodeBase,
                  idx[cur], box, nodes, splits);
        nodeBase += nodesHere;
        if (nSplits == 0) break;
        const uint32_t words = nSplits * static_cast<uint32_t>(kAcc);
        RK_LAUNCH("k_bvh_init_acc", k_bvh_init_acc, dim3(groups(words)), dim3(256), 0, s, words, acc);
        RK_LAUNCH("k_bvh_rank", k_bvh_rank, dim3(groups(n)), dim3(256), 0, s, n, segOf[cur], segs[cur], flags64, rank);
        RK_LAUNCH("k_bvh_bounds", k_bvh_bounds, dim3(groups(n)), dim3(256), 0, s, n, idx[cur], rank, box, acc);
        RK_LAUNCH("k_bvh_params", k_bvh_params, dim3(groups(nSplits)), dim3(256), 0, s, nSplits, idx[cur], box, acc,
                  params);
        RK_LAUNCH("k_bvh_bin", k_bvh_bin, dim3(groups(n)), dim3(256), 0, s, n, idx[cur], rank, box, params, bucket, acc);
        RK_LAUNCH("k_bvh_choose", k_bvh_choose, dim3(groups(nSplits)), dim3(256), 0, s, nSplits, idx[cur], box, params,
                  acc, splits, nodes, segs[cur ^ 1]);
        RK_LAUNCH("k_bvh_halves", k_bvh_halves, dim3(groups(n)), dim3(256), 0, s, n, idx[cur], rank, box, splits, acc);
        RK_LAUNCH("k_bvh_halves_write", k_bvh_halves_write, dim3(groups(nSplits)), dim3(256), 0, s, nSplits, idx[cur],
                  box, splits, acc, nodes);
        RK_LAUNCH("k_bvh_left_flags", k_bvh_left_flags, dim3(groups(n)), dim3(256), 0, s, n, rank, splits, bucket, flag);
        if (int r = hip(rscan::exclusive_scan<uint32_t>(flag, flag, n, temp32, nullptr, s), "build_bvh: scan")) return r;
        RK_LAUNCH("k_bvh_chain", k_bvh_chain, dim3(groups(n)), dim3(256), 0, s, n, rank, splits, bucket, flag, ptr[0]);
        int pc = 0;
        for (uint32_t span = 1; span < host->maxLen; span *= 2) {  // ceil(log2(longest range)) rounds
            RK_LAUNCH("k_bvh_jump", k_bvh_jump, dim3(groups(n)), dim3(256), 0, s, n, ptr[pc], ptr[pc ^ 1]);
            pc ^= 1;
        }
        RK_LAUNCH("k_bvh_scatter", k_bvh_scatter, dim3(groups(n)), dim3(256), 0, s, n, rank, splits, bucket, flag,
                  ptr[pc], idx[cur], idx[cur ^ 1], segOf[cur ^ 1]);
        if (int r = hip(hipGetLastError(), "build_bvh: launch")) return r;
        S = 2 * nSplits;
        cur ^= 1;
    }
    if (int r = hip(hipStreamSynchronize(s), "build_bvh: finish")) return r;
    if (nodeBase != n - 1) return fail(ctx, RESTIR_E_BAD_TREE, "build_bvh: %u nodes for %u triangles", nodeBase, n);
    if (depth > 64)
        return fail(ctx, RESTIR_E_BAD_TREE, "build_bvh: tree depth %u exceeds the 64-entry traversal stack", depth);
    if (depthOut) *depthOut = depth;
    return bind_tree(ctx, nodes, n - 1, tris, depth, s);
}
