// host_build.cpp — host-side builders of the scene tables the ReSTIR passes consume, producing the
// reference's exact buffer contents:
//   restir_build_aabb_tree   <- AabbTree::build      (src/aabbTreeBuilder.cpp:52-214)
//   restir_build_alias_table <- createAliasTable      (src/misc.cpp:418-497)
//   restir_build_light_clusters (k-means light clusters of the learned proposal; its nearest-
//   centroid step also runs on the GPU, cluster_build.hip)
// Single-threaded like the reference; float arithmetic in the reference's order (no contraction).
#include <cfloat>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <deque>
#include <queue>
#include <vector>

#include "../../include/restir_abi.h"
#include "clusters.h"

namespace {

// nvmath::nv_min / nv_max (thirdparty/nvmath/nvmath.inl:2414-2420)
inline float nvMin(float a, float b) { return (a < b) ? a : b; }
inline float nvMax(float a, float b) { return (a > b) ? a : b; }

struct Box {
    float lo[3], hi[3];
};
inline float halfArea(const Box &b) {  // surfaceAreaHeuristic, aabbTreeBuilder.cpp:15-18
    const float sx = b.hi[0] - b.lo[0], sy = b.hi[1] - b.lo[1], sz = b.hi[2] - b.lo[2];
    return sx * sy + sx * sz + sy * sz;
}
inline void grow(Box &a, const Box &b) {
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = nvMin(a.lo[k], b.lo[k]);
        a.hi[k] = nvMax(a.hi[k], b.hi[k]);
    }
}

struct Prim {
    float centroid[3];
    Box box;
    int32_t tri;
    uint32_t bucket;
};

struct Bin {  // Bucket, aabbTreeBuilder.cpp:33-50
    Box box{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    size_t count = 0;
    float cost() const { return static_cast<float>(count) * halfArea(box); }
    static Bin merged(Bin a, const Bin &b) {
        a.count += b.count;
        grow(a, b);
        return a;
    }
    static void grow(Bin &a, const Bin &b) { ::grow(a.box, b.box); }
};

void writeBox(float dst[4], const float src[3]) {
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
    dst[3] = 1.0f;  // nvmath vec4(vec3) sets w = 1
}

struct Task {
    int32_t *link;  // where the child reference is written
    size_t first, last;
    uint32_t depth;
};

}  // namespace

extern "C" int restir_build_aabb_tree(const float *triXyz, uint32_t nTris,
                                      restir_aabb_tree_node *nodes, restir_triangle *tris,
                                      uint32_t *depthOut) {
    if (!triXyz || !nodes || !tris || nTris < 2) return RESTIR_E_INVALID_ARGUMENT;
    constexpr size_t kBuckets = 12;
    std::vector<Prim> prims(nTris);
    for (uint32_t i = 0; i < nTris; ++i) {
        const float *v = triXyz + 9 * static_cast<size_t>(i);
        restir_triangle &t = tris[i];
        for (int k = 0; k < 3; ++k) {
            t.p1[k] = v[k];
            t.p2[k] = v[3 + k];
            t.p3[k] = v[6 + k];
        }
        t.p1[3] = t.p2[3] = t.p3[3] = 1.0f;
        Prim &p = prims[i];
        for (int k = 0; k < 3; ++k) {  // aabbForTriangle
            float lo = v[k], hi = v[k];
            lo = nvMin(lo, v[3 + k]);
            hi = nvMax(hi, v[3 + k]);
            lo = nvMin(lo, v[6 + k]);
            hi = nvMax(hi, v[6 + k]);
            p.box.lo[k] = lo;
            p.box.hi[k] = hi;
            p.centroid[k] = 0.5f * (lo + hi);
        }
        p.tri = static_cast<int32_t>(i);
        p.bucket = 0;
    }
    std::memset(nodes, 0, sizeof(restir_aabb_tree_node) * (nTris - 1));

    int32_t next = 0;
    int32_t rootLink = -1;
    uint32_t maxDepth = 0;
    std::deque<Task> fifo;  // breadth-first node numbering
    fifo.push_back(Task{&rootLink, 0, nTris, 1});
    while (!fifo.empty()) {
        const Task task = fifo.front();
        fifo.pop_front();
        const size_t n = task.last - task.first;
        if (n == 1) {
            *task.link = ~prims[task.first].tri;
            continue;
        }
        const int32_t id = next++;
        *task.link = id;
        if (task.depth > maxDepth) maxDepth = task.depth;
        restir_aabb_tree_node &node = nodes[id];
        if (n == 2) {
            const Prim &a = prims[task.first], &b = prims[task.first + 1];
            node.leftChild = ~a.tri;
            node.rightChild = ~b.tri;
            writeBox(node.leftAabbMin, a.box.lo);
            writeBox(node.leftAabbMax, a.box.hi);
            writeBox(node.rightAabbMin, b.box.lo);
            writeBox(node.rightAabbMax, b.box.hi);
            continue;
        }
        // bounds of centroids and of boxes over the range
        Box cb{{prims[task.first].centroid[0], prims[task.first].centroid[1], prims[task.first].centroid[2]},
               {prims[task.first].centroid[0], prims[task.first].centroid[1], prims[task.first].centroid[2]}};
        Box bb = prims[task.first].box;
        for (size_t i = task.first + 1; i < task.last; ++i) {
            Box c{{prims[i].centroid[0], prims[i].centroid[1], prims[i].centroid[2]},
                  {prims[i].centroid[0], prims[i].centroid[1], prims[i].centroid[2]}};
            grow(cb, c);
            grow(bb, prims[i].box);
        }
        const float parentArea = halfArea(bb);
        float extent[3];
        for (int k = 0; k < 3; ++k) extent[k] = cb.hi[k] - cb.lo[k];
        int axis = extent[0] > extent[1] ? 0 : 1;
        if (extent[2] > extent[axis]) axis = 2;

        Bin bins[kBuckets];
        const float width = extent[axis] / static_cast<float>(kBuckets);
        for (size_t i = task.first; i < task.last; ++i) {
            float f = (prims[i].centroid[axis] - cb.lo[axis]) / width;
            f = (f < 0.5f) ? 0.5f : f;  // nv_clamp(f, 0.5, numBuckets - 0.5)
            f = (f > static_cast<float>(kBuckets) - 0.5f) ? static_cast<float>(kBuckets) - 0.5f : f;
            // coincident centroids give NaN (undefined in the reference); use bucket 0
            const uint32_t bi = (f == f) ? static_cast<uint32_t>(f) : 0u;
            prims[i].bucket = bi;
            Bin one;
            one.box = prims[i].box;
            Bin::grow(bins[bi], one);
            bins[bi].count++;
        }
        // suffix unions: right[s] = bins s+1..11, merged from the top bucket down
        Bin right[kBuckets - 1];
        {
            Bin acc = bins[kBuckets - 1];
            for (size_t s = kBuckets - 1; s > 0;) {
                right[--s] = acc;
                acc = Bin::merged(acc, bins[s]);
            }
        }
        size_t split = 0;
        Box lbox{{0, 0, 0}, {0, 0, 0}}, rbox{{0, 0, 0}, {0, 0, 0}};
        {
            float best = FLT_MAX;
            Bin left;
            for (size_t s = 0; s < kBuckets - 1; ++s) {
                left = Bin::merged(left, bins[s]);
                const float cost = 0.125f + (left.cost() + right[s].cost()) / parentArea;
                if (cost < best) {
                    best = cost;
                    split = s;
                    lbox = left.box;
                    rbox = right[s].box;
                }
            }
        }
        size_t mid = task.first;
        for (size_t i = task.first; i < task.last; ++i) {
            if (prims[i].bucket <= split) {
                std::swap(prims[i], prims[mid]);
                ++mid;
            }
        }
        if (mid == task.first || mid == task.last) {  // degenerate split: halve the range
            mid = (task.first + task.last) / 2;
            lbox = prims[task.first].box;
            for (size_t i = task.first + 1; i < mid; ++i) grow(lbox, prims[i].box);
            rbox = prims[mid].box;
            for (size_t i = mid; i < task.last; ++i) grow(rbox, prims[i].box);
        }
        writeBox(node.leftAabbMin, lbox.lo);
        writeBox(node.leftAabbMax, lbox.hi);
        writeBox(node.rightAabbMin, rbox.lo);
        writeBox(node.rightAabbMax, rbox.hi);
        fifo.push_back(Task{&node.leftChild, task.first, mid, task.depth + 1});
        fifo.push_back(Task{&node.rightChild, mid, task.last, task.depth + 1});
    }
    if (depthOut) *depthOut = maxDepth;
    return (rootLink == 0 && next == static_cast<int32_t>(nTris - 1)) ? RESTIR_OK : RESTIR_E_BAD_TREE;
}

namespace {

// Vose's method with FIFO queues over `n` weights (createAliasTable, misc.cpp:418-497); `alias`
// entries are local indices 0..n-1.
void vose(const float *weights, uint32_t n, restir_alias_table_column *cols) {
    std::vector<float> scaled(n);
    float total = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        total += weights[i];
        scaled[i] = weights[i];
    }
    std::queue<uint32_t> over, under;
    for (uint32_t i = 0; i < n; ++i) {
        cols[i].prob = 0.0f;
        cols[i].alias = -1;
        cols[i].oriProb = scaled[i] / total;
        cols[i].aliasOriProb = 0.0f;
        scaled[i] = static_cast<float>(n) * scaled[i] / total;
        (scaled[i] >= 1.0f ? over : under).push(i);
    }
    while (!over.empty() && !under.empty()) {
        const uint32_t g = over.front();
        over.pop();
        const uint32_t l = under.front();
        under.pop();
        cols[l].prob = scaled[l];
        cols[l].alias = static_cast<int32_t>(g);
        scaled[g] = (scaled[g] + scaled[l]) - 1.0f;
        (scaled[g] < 1.0f ? under : over).push(g);
    }
    for (; !over.empty(); over.pop()) {
        cols[over.front()].prob = 1.0f;
        cols[over.front()].alias = static_cast<int32_t>(over.front());
    }
    for (; !under.empty(); under.pop()) {
        cols[under.front()].prob = 1.0f;
        cols[under.front()].alias = static_cast<int32_t>(under.front());
    }
    for (uint32_t i = 0; i < n; ++i) cols[i].aliasOriProb = cols[cols[i].alias].oriProb;
}

// PCG32 (the generator of rand.glsl:12-28), used to seed k-means reproducibly.
struct Pcg32 {
    uint64_t state, inc;
    Pcg32(uint64_t seed, uint64_t seq) : state(0), inc((seq << 1u) | 1u) {
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
};

inline uint32_t nearest(const float *c, uint32_t k, float x, float y, float z) {
    uint32_t best = 0;
    float bestD = 0.0f;
    for (uint32_t j = 0; j < k; ++j) {
        const float dx = x - c[3 * j], dy = y - c[3 * j + 1], dz = z - c[3 * j + 2];
        const float d = (dx * dx + dy * dy) + dz * dz;
        if (j == 0 || d < bestD) {
            bestD = d;
            best = j;
        }
    }
    return best;
}

}  // namespace

namespace rki {

int build_light_clusters(const restir_point_light *points, uint32_t n, uint32_t k, uint64_t seed,
                         uint32_t iterations, const NearestStep &nearestStep, int32_t *lightCluster,
                         float *lightPWithin, int32_t *clusterBegin, int32_t *clusterCount,
                         int32_t *clusterLights, restir_alias_table_column *clusterAlias, float *centroidsOut) {
    if (!points || !lightCluster || !lightPWithin || !clusterBegin || !clusterCount || !clusterLights ||
        !clusterAlias || k == 0 || k > n)
        return RESTIR_E_INVALID_ARGUMENT;
    // initial centroids: a seeded partial Fisher-Yates draw of k distinct lights
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    Pcg32 rng(seed, 0x6b6d65616e73ull);
    std::vector<float> c(3 * static_cast<size_t>(k));
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t j = i + rng.next() % (n - i);
        std::swap(perm[i], perm[j]);
        for (int d = 0; d < 3; ++d) c[3 * i + d] = points[perm[i]].pos[d];
    }
    std::vector<uint32_t> assign(n);
    std::vector<double> sum(3 * static_cast<size_t>(k));
    std::vector<uint32_t> cnt(k);
    for (uint32_t it = 0; it <= iterations; ++it) {
        if (int r = nearestStep(c.data(), assign.data())) return r;
        if (it == iterations) break;  // the final assignment uses the final centroids
        std::fill(sum.begin(), sum.end(), 0.0);
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (uint32_t l = 0; l < n; ++l) {
            for (int d = 0; d < 3; ++d) sum[3 * assign[l] + d] += static_cast<double>(points[l].pos[d]);
            ++cnt[assign[l]];
        }
        for (uint32_t j = 0; j < k; ++j)
            if (cnt[j] > 0)
                for (int d = 0; d < 3; ++d) c[3 * j + d] = static_cast<float>(sum[3 * j + d] / cnt[j]);
    }
    // lights grouped by cluster, in index order within a cluster
    std::fill(cnt.begin(), cnt.end(), 0u);
    for (uint32_t l = 0; l < n; ++l) ++cnt[assign[l]];
    int32_t off = 0;
    for (uint32_t j = 0; j < k; ++j) {
        clusterBegin[j] = off;
        clusterCount[j] = static_cast<int32_t>(cnt[j]);
        off += static_cast<int32_t>(cnt[j]);
    }
    std::vector<int32_t> fill(clusterBegin, clusterBegin + k);
    for (uint32_t l = 0; l < n; ++l) {
        lightCluster[l] = static_cast<int32_t>(assign[l]);
        clusterLights[fill[assign[l]]++] = static_cast<int32_t>(l);
    }
    // per-cluster alias tables over luminance; oriProb is the within-cluster probability
    std::vector<float> w;
    for (uint32_t j = 0; j < k; ++j) {
        const int32_t b = clusterBegin[j], m = clusterCount[j];
        if (m == 0) continue;
        w.resize(m);
        for (int32_t i = 0; i < m; ++i) w[i] = points[clusterLights[b + i]].color_luminance[3];
        vose(w.data(), static_cast<uint32_t>(m), clusterAlias + b);
        for (int32_t i = 0; i < m; ++i) lightPWithin[clusterLights[b + i]] = clusterAlias[b + i].oriProb;
    }
    if (centroidsOut) std::copy(c.begin(), c.end(), centroidsOut);
    return RESTIR_OK;
}

}  // namespace rki

extern "C" int restir_build_light_clusters(const restir_point_light *points, uint32_t n, uint32_t k,
                                           uint64_t seed, uint32_t iterations, int32_t *lightCluster,
                                           float *lightPWithin, int32_t *clusterBegin,
                                           int32_t *clusterCount, int32_t *clusterLights,
                                           restir_alias_table_column *clusterAlias,
                                           float *centroidsOut) {
    const rki::NearestStep host = [&](const float *c, uint32_t *assign) {
        for (uint32_t l = 0; l < n; ++l) assign[l] = nearest(c, k, points[l].pos[0], points[l].pos[1], points[l].pos[2]);
        return 0;
    };
    return rki::build_light_clusters(points, n, k, seed, iterations, host, lightCluster, lightPWithin, clusterBegin,
                                     clusterCount, clusterLights, clusterAlias, centroidsOut);
}

extern "C" int restir_build_alias_table(const restir_point_light *points, uint32_t nPoint,
                                        const restir_tri_light *triLights, uint32_t nTri,
                                        restir_alias_table_column *cols) {
    const uint32_t n = nPoint > 0 ? nPoint : nTri;
    if (n == 0) return RESTIR_OK;
    if (!cols || (nPoint > 0 && !points) || (nPoint == 0 && !triLights)) return RESTIR_E_INVALID_ARGUMENT;
    // point lights weigh by luminance, triangle lights by luminance * area (misc.cpp:426-442)
    std::vector<float> power(n);
    for (uint32_t i = 0; i < n; ++i)
        power[i] = nPoint > 0 ? points[i].color_luminance[3]
                              : triLights[i].emission_luminance[3] * triLights[i].normalArea[3];
    vose(power.data(), n, cols);
    return RESTIR_OK;
}

// Extension (no reference counterpart): one table over the point lights followed by the triangle
// lights, with createAliasTable's weights and construction (misc.cpp:418-497).
extern "C" int restir_build_alias_table_mixed(const restir_point_light *points, uint32_t nPoint,
                                              const restir_tri_light *triLights, uint32_t nTri,
                                              restir_alias_table_column *cols) {
    const uint64_t n = static_cast<uint64_t>(nPoint) + nTri;
    if (n == 0) return RESTIR_OK;
    if (n > (1ull << 30)) return RESTIR_E_INVALID_ARGUMENT;
    if (!cols || (nPoint > 0 && !points) || (nTri > 0 && !triLights)) return RESTIR_E_INVALID_ARGUMENT;
    std::vector<float> power(static_cast<size_t>(n));
    for (uint32_t i = 0; i < nPoint; ++i) power[i] = points[i].color_luminance[3];
    for (uint32_t i = 0; i < nTri; ++i)
        power[nPoint + i] = triLights[i].emission_luminance[3] * triLights[i].normalArea[3];
    vose(power.data(), static_cast<uint32_t>(n), cols);
    return RESTIR_OK;
}
