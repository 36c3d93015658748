// passes.hip — gfx950 kernels of the ReSTIR pass structure:
//   k_pass_a   initial RIS + visibility reuse + temporal reuse   (restirOmni.glsl:86-212)
//   k_pass_b   biased spatial reuse                              (spatialReuse.comp:30-85)
//   k_pass_c   unbiased spatial reuse                            (unbiasedReuse.glsl:50-185)
//   k_shade    final shading                                     (lighting.frag:49-71, 103)
//   k_trace    batch shadow rays                                 (visibilityTest.glsl, softwareRaytracing.glsl)
//   k_primary  primary-ray G-buffer (closest hit)                — synthetic-input generator
//
// One lane per pixel. Workgroups cover a tx*ty pixel tile; tiles are handed to the 8 XCDs in
// contiguous runs so that neighbouring tiles (which gather each other's reservoirs) share an L2.
// G-buffer planes are read as SoA float planes (one coalesced dword per lane per plane);
// reservoirs keep the reference's 64-byte AoS record and move as four 16-byte vectors.
#include <cstring>
#include <type_traits>
#include <vector>

#include "carve.h"
#include "kernels.h"
#include "scan.h"

namespace rk {

using namespace rt;

const Tile kTiles[kNumTiles] = {{16, 16}, {64, 1}, {8, 8}, {16, 8}, {32, 8}};

int tile_index(uint32_t tx, uint32_t ty) {
    for (int i = 0; i < kNumTiles; ++i)
        if (kTiles[i].tx == tx && kTiles[i].ty == ty) return i;
    return -1;
}

// Calls f with the compiled traversal-stack depth that covers `stack` entries (24, 32 or 64
// dwords per lane of LDS: the smaller the stack, the more workgroups share a CU).
template <class F>
static hipError_t by_stack(int stack, F &&f) {
    if (stack <= 24) return f(std::integral_constant<int, 24>{});
    if (stack <= 32) return f(std::integral_constant<int, 32>{});
    return f(std::integral_constant<int, 64>{});
}

// Hardware places workgroup b on XCD b % 8. Remap so XCD k processes tiles
// [base_k, base_k + count_k) in raster order.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
    constexpr uint32_t kX = 8;
    if (n < kX) return b;
    const uint32_t per = n / kX, rem = n % kX;
    const uint32_t x = b % kX, slot = b / kX;
    return x * per + (x < rem ? x : rem) + slot;
}

// Tile t of the blocked order: bands of kBandRows tile rows, each walked in column blocks of
// kBlockCols tiles (the last band and block may be smaller). An XCD's run of consecutive tiles
// is then a compact screen region, so the reservoirs that neighbouring pixels gather (up to
// 30 px away) are shared in its L2 instead of spanning the full screen width.
constexpr uint32_t kBandRows = 16, kBlockCols = 16;
__device__ __forceinline__ void blocked_tile(uint32_t t, uint32_t tilesX, uint32_t tilesY,
                                             uint32_t &tx, uint32_t &ty) {
    const uint32_t band = t / (kBandRows * tilesX), r = t % (kBandRows * tilesX);
    const uint32_t h = min(kBandRows, tilesY - band * kBandRows);
    const uint32_t cb = r / (kBlockCols * h), rr = r % (kBlockCols * h);
    const uint32_t w = min(kBlockCols, tilesX - cb * kBlockCols);
    tx = cb * kBlockCols + rr % w;
    ty = band * kBandRows + rr / w;
}

template <int TX, int TY>
__device__ __forceinline__ void tile_pixel(uint32_t tilesX, uint32_t nTiles, uint32_t &px,
                                           uint32_t &py) {
    const uint32_t t = xcd_tile(blockIdx.x, nTiles);
    uint32_t tx, ty;
    blocked_tile(t, tilesX, nTiles / tilesX, tx, ty);
    const uint32_t tid = threadIdx.x;
    px = tx * TX + tid % TX;
    py = ty * TY + tid / TX;
}

__device__ __forceinline__ f3 plane3(const float *__restrict__ p, size_t i, uint64_t P) {
    return f3{p[i], p[P + i], p[2 * P + i]};
}

__device__ __forceinline__ Res load_res(const restir_reservoir *__restrict__ base, size_t i) {
    const float4 *q = reinterpret_cast<const float4 *>(base + i);
    const float4 a = q[0], b = q[1], c = q[2], d = q[3];
    Res r;
    r.pos = f3{a.x, a.y, a.z};
    r.emLum = a.w;
    r.normal = b;
    r.lightIndex = __float_as_int(c.x);
    r.pHat = c.y;
    r.sumW = c.z;
    r.w = c.w;
    r.M = __float_as_uint(d.x);
    return r;
}

__device__ __forceinline__ void store_res(restir_reservoir *__restrict__ base, size_t i,
                                          const Res &r) {
    float4 *q = reinterpret_cast<float4 *>(base + i);
    q[0] = make_float4(r.pos.x, r.pos.y, r.pos.z, r.emLum);
    q[1] = r.normal;
    q[2] = make_float4(__int_as_float(r.lightIndex), r.pHat, r.sumW, r.w);
    q[3] = make_float4(__uint_as_float(r.M), 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ float phat_of(const Surface &s, const Res &r) {
    return eval_phat(s, r.pos, f3{r.normal.x, r.normal.y, r.normal.z}, r.normal.w > 0.5f, r.emLum);
}

// ------------------------------------------------------------------------------- Pass A ---
struct PassAArgs {
    restir_uniforms u;
    SceneView sc;
    GView cur, prev;
    const restir_reservoir *prevRes;
    restir_reservoir *out;
    uint32_t tilesX, nTiles;
    // learned proposal: candidate j of pixel p is (candLight, candQ)[j * npx + p]; null = the
    // reference's alias-table candidates
    const int32_t *candLight;
    const float *candQ;
};

// Pass A in pieces (restirOmni.glsl:86-212), shared by the inline kernel and the queued schedule.
struct PassASurface {
    size_t idx;
    f3 pos, n, alb;
    Surface s;
};
__device__ __forceinline__ PassASurface pass_a_surface(const PassAArgs &a, uint32_t x, uint32_t y) {
    PassASurface p;
    p.idx = static_cast<size_t>(y) * a.u.screenSize[0] + x;
    const GView &g = a.cur;
    p.pos = plane3(g.pos, p.idx, g.P);
    p.n = load_normal(g, p.idx);
    p.alb = plane3(g.alb, p.idx, g.P);
    const float rough = g.mat[p.idx], metal = g.mat[g.P + p.idx];
    const f3 cam{a.u.cameraPos[0], a.u.cameraPos[1], a.u.cameraPos[2]};
    p.s = make_surface(p.pos, p.n, cam, luminance(p.alb.x, p.alb.y, p.alb.z), rough, metal);
    return p;
}
__device__ __forceinline__ Pcg pass_a_rng(const PassAArgs &a, uint32_t x, uint32_t y) {
    return pcg_seed(static_cast<uint64_t>(a.u.frame), static_cast<uint64_t>(y * 10007u + x));
}

// One light candidate of Pass A (restirOmni.glsl:113-145) with source probability prob: point
// light li (`point`: its position and luminance, point_light()), or (tri) a point on triangle
// light li picked with the draws ra, rb
// (pickPointOnTriangle, restirOmni.glsl:68-71), whose probability per unit area seen from pos is
// prob / (|cos| area); u is the reservoir update's draw. Only the choice of the sample point
// depends on the kind; the target function and the reservoir update run once for all lanes of a
// wave whatever kinds its lanes drew. `valid` false: an empty candidate (p_hat 0, probability 1)
// that still counts in M and draws its update number.
__device__ __forceinline__ void add_candidate(Res &res, const Surface &s, f3 pos, const SceneView &sc, bool valid,
                                              bool tri, int32_t li, float prob, float4 point, float ra, float rb,
                                              float u, bool areaPdf) {
    f3 L{0.0f, 0.0f, 0.0f}, nl{0.0f, 0.0f, 0.0f};
    float lum = 0.0f;
    int32_t index = 0;
    if (valid && !tri) {
        L = xyz(point);
        lum = point.w;
        index = li;
    } else if (valid) {
        const restir_tri_light &tl = sc.triLights[li];
        const f3 p1 = xyz(*reinterpret_cast<const float4 *>(tl.p1));
        const f3 p2 = xyz(*reinterpret_cast<const float4 *>(tl.p2));
        const f3 p3 = xyz(*reinterpret_cast<const float4 *>(tl.p3));
        const float4 na = *reinterpret_cast<const float4 *>(tl.normalArea);
        lum = tl.emission_luminance[3];
        const float sa = sqrtf(ra);
        L = ((1.0f - sa) * p1 + (sa * (1.0f - rb)) * p2) + (rb * sa) * p3;
        const f3 wi = normalize(pos - L);
        nl = xyz(na);
        if (areaPdf) prob /= na.w;  // uniform by area (RESTIR_STANDARD_WEIGHTS_FLAG)
        else prob /= fabsf(dot(wi, nl)) * na.w;
        index = -1 - li;
    }
    const bool useLN = valid && tri;
    const float ph = valid ? eval_phat(s, L, nl, useLN, lum) : 0.0f;
    res_add_sample_u(res, L, make_float4(nl.x, nl.y, nl.z, useLN ? 1.0f : 0.0f), lum, index, ph,
                     valid ? prob : 1.0f, u);
}

// Point light li's position and luminance (pointLight, light.glsl).
__device__ __forceinline__ float4 point_light(const SceneView &sc, int32_t li) {
    const float4 p = *reinterpret_cast<const float4 *>(sc.points[li].pos);
    return make_float4(p.x, p.y, p.z, sc.points[li].color_luminance[3]);
}

// aliasTableSample's column, restirOmni.glsl:73-83
__device__ __forceinline__ int32_t alias_column(const SceneView &sc, float r1) {
    const int32_t col = static_cast<int32_t>(static_cast<float>(sc.nAlias) * r1);
    return col < sc.nAlias - 1 ? col : sc.nAlias - 1;
}

// Initial RIS over M candidates (alias table or the learned proposal's list).
__device__ __forceinline__ void pass_a_ris(const PassAArgs &a, const PassASurface &ps, Res &res, Pcg &rng) {
    const uint32_t W = a.u.screenSize[0], H = a.u.screenSize[1];
    const size_t idx = ps.idx;
    const f3 pos = ps.pos, n = ps.n;
    const Surface &s = ps.s;
    res = res_zero();
    const SceneView &sc = a.sc;
    const bool standard = (a.u.flags & RESTIR_STANDARD_WEIGHTS_FLAG) != 0;
    const uint32_t M = a.u.initialLightSampleCount;
    if (dot(n, n) != 0.0f && a.candLight) {
        // RIS over the learned proposal's candidates: weight p_hat / q with the proposal's own
        // density q; a candidate from an empty cluster (l < 0) has weight 0 but still counts in M.
        // Candidates index the lights the alias table indexes (the union: point lights first).
        const size_t npx = static_cast<size_t>(W) * H;
        for (uint32_t i = 0; i < M; ++i) {
            const int32_t li = a.candLight[i * npx + idx];
            const float q = a.candQ[i * npx + idx];
            const bool valid = li >= 0 && li < sc.nLights && q > 0.0f;
            const bool tri = li >= sc.nPoints;
            float ra = 0.0f, rb = 0.0f;
            if (valid && tri) {
                ra = pcg_float(rng);
                rb = pcg_float(rng);
            }
            const float u = pcg_float(rng);
            const float4 point = valid && !tri ? point_light(sc, li) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            add_candidate(res, s, pos, sc, valid, tri, tri ? li - sc.nPoints : li, q, point, ra, rb, u, standard);
        }
    } else if (dot(n, n) != 0.0f && M > 0) {
        // A candidate draws r1, r2 (the alias column and its coin), then a triangle light's two
        // point draws, then the reservoir update's: 3 draws, or 5 after a triangle light. The
        // candidates run as a pipeline: while candidate i is evaluated, candidate i + 1 is already
        // resolved (its column was loaded during candidate i - 1) with its point light's record in
        // flight, and candidate i + 2's column is in flight, each from its exact first state
        // (jumping 3 or 5 steps ahead once the kind before it is known). Same draws in the same
        // order as drawing them one by one.
        const uint64_t inc = rng.inc;
        const uint64_t c3 = inc * pcg_mul_sum(3), c5 = inc * pcg_mul_sum(5);
        // the state after candidate i's draws, from its first state and kind
        const auto after = [&](uint64_t first, bool tri) {
            return tri ? first * pcg_mul_pow(5) + c5 : first * pcg_mul_pow(3) + c3;
        };
        // aliasTableSample (restirOmni.glsl:73-83) from column col (record cv) and the coin r2;
        // restirOmni.glsl:117: point lights if there are any, else triangle lights; a union table
        // names point lights first
        struct Cand {
            int32_t li;  // index among its kind's lights
            float prob;
            bool tri;
            float4 point;  // a point light's record
        };
        const auto resolve = [&](int32_t col, float4 cv, float r2) {
            Cand k;
            int32_t li;
            if (cv.x > r2) {
                li = col;
                k.prob = cv.z;
            } else {
                li = __float_as_int(cv.y);
                k.prob = cv.w;
            }
            k.tri = sc.mixed ? li >= sc.nPoints : sc.nPoints == 0;
            k.li = sc.mixed && k.tri ? li - sc.nPoints : li;
            k.point = k.tri ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : point_light(sc, k.li);
            return k;
        };
        uint64_t st = rng.state;  // candidate i's first state
        int32_t col = alias_column(sc, pcg_float_at(st));
        Cand cur = resolve(col, sc.alias[col], pcg_float_at(st * kPcgMul + inc));
        uint64_t stN = after(st, cur.tri);  // candidate i + 1's first state
        int32_t colN = alias_column(sc, pcg_float_at(stN));
        float4 cN = sc.alias[colN];  // candidate i + 1's column
        for (uint32_t i = 0; i < M; ++i) {
            Cand nxt = cur;
            uint64_t stNN = stN;
            int32_t colNN = colN;
            float4 cNN = cN;
            if (i + 1 < M) {
                nxt = resolve(colN, cN, pcg_float_at(stN * kPcgMul + inc));
                stNN = after(stN, nxt.tri);
                colNN = alias_column(sc, pcg_float_at(stNN));
                cNN = sc.alias[colNN];
            }
            const uint64_t s2 = st * pcg_mul_pow(2) + inc * pcg_mul_sum(2), s3 = s2 * kPcgMul + inc;
            float ra = 0.0f, rb = 0.0f, u;
            if (cur.tri) {
                ra = pcg_float_at(s2);
                rb = pcg_float_at(s3);
                u = pcg_float_at(s3 * kPcgMul + inc);
            } else {
                u = pcg_float_at(s2);
            }
            add_candidate(res, s, pos, sc, true, cur.tri, cur.li, cur.prob, cur.point, ra, rb, u, standard);
            st = stN;
            cur = nxt;
            stN = stNN;
            colN = colNN;
            cN = cNN;
        }
        rng.state = st;
    }
    // the standard RIS contribution weight instead of the selection-time one
    if (standard && res.w > 0.0f) res.w = res.sumW / (static_cast<float>(res.M) * res.pHat);
}

// Temporal reuse of the previous frame's reservoir at the reprojected pixel.
__device__ __forceinline__ void pass_a_temporal(const PassAArgs &a, const PassASurface &ps, Res &res, Pcg &rng) {
    const uint32_t W = a.u.screenSize[0], H = a.u.screenSize[1];
    const f3 pos = ps.pos, n = ps.n, alb = ps.alb;
    const Surface &s = ps.s;
    // Temporal reuse, restirOmni.glsl:163-209 (COMPARE_DEPTH off: world-position test).
    if (a.u.flags & RESTIR_TEMPORAL_REUSE_FLAG) {
        const float *m = a.u.prevFrameProjectionViewMatrix;
        float pp[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
            pp[r] = m[r] * pos.x + m[4 + r] * pos.y + m[8 + r] * pos.z + m[12 + r] * 1.0f;
        pp[0] /= pp[3];
        pp[1] /= pp[3];
        pp[0] = (pp[0] + 1.0f) * 0.5f * static_cast<float>(W);
        pp[1] = (pp[1] + 1.0f) * 0.5f * static_cast<float>(H);
        if (pp[0] > 0.0f && pp[1] > 0.0f && pp[0] < static_cast<float>(W) &&
            pp[1] < static_cast<float>(H)) {
            const uint32_t px = static_cast<uint32_t>(static_cast<int32_t>(pp[0]));
            const uint32_t py = static_cast<uint32_t>(static_cast<int32_t>(pp[1]));
            const size_t pi = static_cast<size_t>(py) * W + px;
            const GView &pg = a.prev;
            const f3 pd = pos - plane3(pg.pos, pi, pg.P);
            if (dot(pd, pd) < 0.01f) {
                const f3 ad = alb - plane3(pg.alb, pi, pg.P);
                if (dot(ad, ad) < 0.01f) {
                    if (dot(n, load_normal(pg, pi)) > 0.5f) {
                        Res pr = load_res(a.prevRes, pi);
                        const uint32_t cap = a.u.temporalSampleCountMultiplier * res.M;
                        pr.M = pr.M < cap ? pr.M : cap;
                        const float ph = phat_of(s, pr);
                        res_combine(res, pr, ph, rng);
                    }
                }
            }
        }
    }
}

// Pass A for pixel (x, y), shadow ray traced inline; `stk` is this lane's column of the
// [level][lane] LDS stack.
template <int TX, int TY, int STACK>
__device__ __forceinline__ void pass_a_px(const PassAArgs &a, uint32_t x, uint32_t y, int32_t *stk) {
    constexpr int kBlock = TX * TY;
    if (x >= a.u.screenSize[0] || y >= a.u.screenSize[1]) return;
    const PassASurface ps = pass_a_surface(a, x, y);
    const size_t idx = ps.idx;
    const f3 pos = ps.pos;
    const SceneView &sc = a.sc;
    Res res;
    Pcg rng = pass_a_rng(a, x, y);
    pass_a_ris(a, ps, res, rng);
    // Visibility reuse, restirOmni.glsl:148-160. A reservoir with sumWeights == 0 already has
    // w == 0 and is unchanged by a shadow hit, so its ray is skipped.
    if ((a.u.flags & RESTIR_VISIBILITY_REUSE_FLAG) && res.sumW != 0.0f) {
        f3 o, d;
        shadow_segment(pos, res.pos, o, d);
        TraceStats st{0, 0};
        if (occluded_w<STACK>(sc.wide, sc.nodes, sc.tris, sc.treeNested != 0, sc.maxCoord, o, d, stk, kBlock, st)) {
            res.w = 0.0f;
            res.sumW = 0.0f;
        }
    }
    pass_a_temporal(a, ps, res, rng);
    store_res(a.out, idx, res);
}

template <int TX, int TY, int STACK>
__global__ __launch_bounds__(TX *TY) void k_pass_a(const PassAArgs a) {
    __shared__ int32_t stack[STACK * TX * TY];
    uint32_t x, y;
    tile_pixel<TX, TY>(a.tilesX, a.nTiles, x, y);
    pass_a_px<TX, TY, STACK>(a, x, y, stack + threadIdx.x);
}

// ------------------------------------------------------------------------------- Pass B ---
struct PassBArgs {
    restir_uniforms u;
    GView g;
    const restir_reservoir *in;
    restir_reservoir *out;
    uint32_t tilesX, nTiles;
    uint32_t seed;
    float cosThr;
};

template <int TX, int TY>
__device__ __forceinline__ void pass_b_px(const PassBArgs &a, uint32_t x, uint32_t y) {
    const uint32_t W = a.u.screenSize[0], H = a.u.screenSize[1];
    if (x >= W || y >= H) return;
    const size_t idx = static_cast<size_t>(y) * W + x;
    const GView &g = a.g;
    const f3 pos = plane3(g.pos, idx, g.P);
    const f3 n = load_normal(g, idx);
    const f3 alb = plane3(g.alb, idx, g.P);
    const float rough = g.mat[idx], metal = g.mat[g.P + idx];
    const float depth = g.depth[idx];
    const f3 cam{a.u.cameraPos[0], a.u.cameraPos[1], a.u.cameraPos[2]};
    const Surface s = make_surface(pos, n, cam, luminance(alb.x, alb.y, alb.z), rough, metal);
    Res res = load_res(a.in, idx);
    Pcg rng = pcg_seed(static_cast<uint64_t>(a.seed), static_cast<uint64_t>(y * 10007u + x));
    const float R = a.u.spatialRadius;
    const float depthTol = a.u.spatialPosThreshold * fabsf(depth);
    for (uint32_t i = 0; i < a.u.spatialNeighbors; ++i) {
        const float angle = pcg_float(rng) * 2.0f * kPi;
        const float radius = sqrtf(pcg_float(rng)) * R;
        float sn, cs;
        sincos_f(angle, sn, cs);
        int32_t nx = static_cast<int32_t>(x) + static_cast<int32_t>(floorf(cs * radius));
        int32_t ny = static_cast<int32_t>(y) + static_cast<int32_t>(floorf(sn * radius));
        nx = nx < 0 ? 0 : (nx > static_cast<int32_t>(W) - 1 ? static_cast<int32_t>(W) - 1 : nx);
        ny = ny < 0 ? 0 : (ny > static_cast<int32_t>(H) - 1 ? static_cast<int32_t>(H) - 1 : ny);
        const size_t ni = static_cast<size_t>(ny) * W + static_cast<size_t>(nx);
        const float nd = g.depth[ni];
        const f3 nn = load_normal(g, ni);
        if (fabsf(nd - depth) > depthTol || dot(nn, n) < a.cosThr) continue;
        const Res rr = load_res(a.in, ni);
        const float ph = phat_of(s, rr);
        res_combine(res, rr, ph, rng);
    }
    store_res(a.out, idx, res);
}

template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) void k_pass_b(const PassBArgs a) {
    uint32_t x, y;
    tile_pixel<TX, TY>(a.tilesX, a.nTiles, x, y);
    pass_b_px<TX, TY>(a, x, y);
}

// ------------------------------------------------------------------------------- Pass C ---
struct PassCArgs {
    restir_uniforms u;
    SceneView sc;
    GView g;
    const restir_reservoir *in;
    restir_reservoir *out;
    uint32_t tilesX, nTiles;
    int k;
    uint32_t seed;
};

// The neighbour merge of unbiasedReuse.glsl:84-124 for pixel (x, y): k random neighbours in a
// disk, each merged into `res` with the target evaluated at this pixel. Records each
// neighbour's pixel index and sample count for the later Z computation.
// `keep(i, ni, M)` records neighbour i's pixel and sample count for the visibility step; *src
// gets the pixel whose input reservoir supplied the merged sample (this pixel when no neighbour's
// was taken), or kNoPixel when that is this pixel's own input and its W is 0.
constexpr uint32_t kNoPixel = 0xffffffffu;
template <typename Keep>
__device__ __forceinline__ void pass_c_merge(const PassCArgs &a, uint32_t x, uint32_t y, size_t idx,
                                             f3 pos, f3 n, Res &res, Keep keep, uint32_t *src = nullptr) {
    const uint32_t W = a.u.screenSize[0], H = a.u.screenSize[1];
    const GView &g = a.g;
    const f3 alb = plane3(g.alb, idx, g.P);
    const float rough = g.mat[idx], metal = g.mat[g.P + idx];
    const f3 cam{a.u.cameraPos[0], a.u.cameraPos[1], a.u.cameraPos[2]};
    const Surface s = make_surface(pos, n, cam, luminance(alb.x, alb.y, alb.z), rough, metal);
    Pcg rng = pcg_seed(static_cast<uint64_t>(a.seed), static_cast<uint64_t>(y * 10007u + x));
    const float R = a.u.spatialRadius;
    const int k = a.k;
    uint32_t from = res.w > 0.0f ? static_cast<uint32_t>(idx) : kNoPixel;
    for (int i = 0; i < k; ++i) {
        const float angle = pcg_float(rng) * 2.0f * kPi;
        const float radius = sqrtf(pcg_float(rng)) * R;
        float sn, cs;
        sincos_f(angle, sn, cs);
        int32_t nx = static_cast<int32_t>(x) + static_cast<int32_t>(roundf(cs * radius));
        int32_t ny = static_cast<int32_t>(y) + static_cast<int32_t>(roundf(sn * radius));
        nx = nx < 0 ? 0 : (nx > static_cast<int32_t>(W) - 1 ? static_cast<int32_t>(W) - 1 : nx);
        ny = ny < 0 ? 0 : (ny > static_cast<int32_t>(H) - 1 ? static_cast<int32_t>(H) - 1 : ny);
        const uint32_t ni = static_cast<uint32_t>(ny) * W + static_cast<uint32_t>(nx);
        const Res rr = load_res(a.in, ni);
        keep(i, ni, rr.M);
        res.M += rr.M;
        const float ph = phat_of(s, rr);
        const float weight = ph * rr.w * static_cast<float>(rr.M);
        if (weight > 0.0f && res_update_u(res, weight, rr.pos, rr.normal, rr.emLum, rr.lightIndex, ph, rr.w,
                                          pcg_float(rng)))
            from = ni;
    }
    if (src) *src = from;
}

// W of a merged reservoir from the count Z of the reservoirs that could have produced its sample
// (unbiasedReuse.glsl:167-181). With RESTIR_STANDARD_WEIGHTS_FLAG an empty reservoir (p_hat 0)
// gets W = 0 instead of 0/0, and sumW is reset to p_hat W M (M the merged count): a later pass
// then weighs this reservoir like every other input, p_hat W M. The reference keeps sumW = p_hat W Z,
// which a second pass weighs against M: each pass after the first is then biased.
__device__ __forceinline__ void finish_w(float &w, float &sumW, float pHat, uint32_t numSamples, uint32_t M,
                                         bool standard) {
    if (numSamples > 0 && (!standard || pHat > 0.0f)) {
        w = sumW / (static_cast<float>(numSamples) * pHat);
    } else {
        w = 0.0f;
        sumW = 0.0f;
    }
    if (standard) sumW = w * static_cast<float>(M) * pHat;
}

template <int TX, int TY, int STACK>
__device__ __forceinline__ void pass_c_px(const PassCArgs &a, uint32_t x, uint32_t y, int32_t *stk) {
    constexpr int kBlock = TX * TY;
    const uint32_t W = a.u.screenSize[0], H = a.u.screenSize[1];
    if (x >= W || y >= H) return;
    const size_t idx = static_cast<size_t>(y) * W + x;
    const GView &g = a.g;
    const f3 pos = plane3(g.pos, idx, g.P);
    const f3 n = load_normal(g, idx);
    Res res = load_res(a.in, idx);
    if (dot(n, n) == 0.0f) {  // background: pass through (reference reads uninitialised fields)
        store_res(a.out, idx, res);
        return;
    }
    const uint32_t origM = res.M;
    uint32_t nIdx[16];
    uint32_t nM[16];
    const int k = a.k;
    pass_c_merge(a, x, y, idx, pos, n, res, [&](int i, uint32_t ni, uint32_t m) {
        nIdx[i] = ni;
        nM[i] = m;
    });
    const bool vis = (a.u.flags & RESTIR_VISIBILITY_REUSE_FLAG) != 0;
    const f3 lp = res.pos;
    uint32_t numSamples = origM;
    // The centre re-test (unbiasedReuse.glsl:157-166) zeroes Z whatever the neighbours give, and
    // the neighbour loop draws no random numbers, so testing the centre first and skipping the
    // neighbour rays when it is shadowed yields the reference's result with fewer rays.
    bool centreShadowed = false;
    if (vis) {
        f3 o, d;
        shadow_segment(pos, lp, o, d);
        TraceStats st{0, 0};
        centreShadowed = occluded_w<STACK>(a.sc.wide, a.sc.nodes, a.sc.tris, a.sc.treeNested != 0, a.sc.maxCoord, o, d, stk, kBlock, st);
    }
    if (centreShadowed) {
        numSamples = 0;
    } else {
        for (int j = 0; j < k; ++j) {
            const uint32_t ni = nIdx[j];
            const f3 nwp = plane3(g.pos, ni, g.P);
            const f3 nn = load_normal(g, ni);
            if (dot(lp - nwp, nn) < 0.0f) continue;
            if (vis) {
                f3 o, d;
                shadow_segment(nwp, lp, o, d);
                TraceStats st{0, 0};
                if (occluded_w<STACK>(a.sc.wide, a.sc.nodes, a.sc.tris, a.sc.treeNested != 0, a.sc.maxCoord, o, d, stk, kBlock, st)) continue;
            }
            numSamples += nM[j];
        }
    }
    finish_w(res.w, res.sumW, res.pHat, numSamples, res.M, (a.u.flags & RESTIR_STANDARD_WEIGHTS_FLAG) != 0);
    store_res(a.out, idx, res);
}

template <int TX, int TY, int STACK>
__global__ __launch_bounds__(TX *TY) void k_pass_c(const PassCArgs a) {
    __shared__ int32_t stack[STACK * TX * TY];
    uint32_t x, y;
    tile_pixel<TX, TY>(a.tilesX, a.nTiles, x, y);
    pass_c_px<TX, TY, STACK>(a, x, y, stack + threadIdx.x);
}

// ------------------------------------------------------------------- Pass C, ray queues ---
// The same pass as k_pass_c, with its shadow rays moved out of the per-pixel lanes into two
// queues traced by one launch of the persistent tracer: (1) merge every pixel, queue its centre
// ray and the rays to its above-horizon neighbours; (2) trace the centre rays, then the
// neighbour rays, skipping those of pixels whose centre ray was found shadowed (they have Z = 0
// whatever their neighbours give); (3) form Z and W. Rays are appended with one atomic per
// workgroup for both queues, slots handed out by prefix sums over the wave and the workgroup, so
// a queue holds only live rays in pixel-tile order.
struct CQueues {
    uint32_t *counter;     // [0] centre rays queued, [1] neighbour rays queued
    float4 *c1;            // centre rays: the pixel's position, .w = pixel | kCentreTag; the other
                           // end is the light sample of `merged[pixel]`
    float4 *n1;            // neighbour rays: the neighbour's position, .w = pixel | slot << 27; the
                           // other end is the light sample of `merged[pixel]`
    const restir_reservoir *merged;  // the pass's output (merged reservoirs)
    uint32_t *origM;       // per pixel
    uint32_t *bits;        // per pixel: above-horizon neighbour mask (bits 0..15), flags below
    uint32_t *occ;         // per pixel: occluded neighbour mask
    uint32_t *nM;          // [slot * npx + pixel]: the neighbour's sample count
    float4 *surf;          // per pixel: {position.xyz, normal.x}, {normal.yz, 0, 0}
    uint32_t npx;
};
constexpr uint32_t kBitBackground = 1u << 29;  // pass-through pixel, already final
constexpr uint32_t kBitDone = 1u << 30;        // no visibility test: W already written
constexpr uint32_t kBitCentreShadowed = 1u << 31;
constexpr uint32_t kCentreTag = 1u << 31;  // a queued ray's tag: the pixel's centre ray

// The set bits of a wave mask below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Reserves room in two queues for every wave of a workgroup of kWaves waves: w0 and w1 are this
// wave's item counts (the same on all its lanes); the queues' counters are the low and high halves
// of *counter. One 64-bit atomic per workgroup (device-wide atomics on one address are served one
// at a time, about 11 ns each on the MI355X, so one per wave would bound the producing kernel);
// returns this wave's first slot in each queue. Every thread of the workgroup must call it (it
// holds two barriers); lds: kWaves + 1 words of the workgroup's shared memory.
template <uint32_t kWaves>
__device__ __forceinline__ uint2 group_reserve(unsigned long long *counter, uint32_t w0, uint32_t w1, uint2 *lds) {
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) lds[wave] = make_uint2(w0, w1);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s0 = 0, s1 = 0;
        for (uint32_t v = 0; v < kWaves; ++v) {
            const uint2 t = lds[v];
            lds[v] = make_uint2(s0, s1);
            s0 += t.x;
            s1 += t.y;
        }
        unsigned long long old = 0;
        if (s0 | s1) old = atomicAdd(counter, (static_cast<unsigned long long>(s1) << 32) | s0);
        lds[kWaves] = make_uint2(static_cast<uint32_t>(old), static_cast<uint32_t>(old >> 32));
    }
    __syncthreads();
    const uint2 off = lds[wave], base = lds[kWaves];
    return make_uint2(base.x + off.x, base.y + off.y);
}

// The lane's rank among the wave's lanes with `emit` set, and their number.
__device__ __forceinline__ uint2 wave_rank(bool emit) {
    const uint64_t m = __ballot(emit);
    return make_uint2(lanes_below(m), static_cast<uint32_t>(__popcll(m)));
}

// Position and normal of every pixel in one 32-byte record: the merge's horizon test and the
// neighbour-ray emission gather them for random neighbours, one sector per neighbour instead of
// one line per plane.
__global__ __launch_bounds__(256) void k_pack_surface(const GView g, uint32_t npx, float4 *surf) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= npx) return;
    const f3 p = plane3(g.pos, i, g.P), n = load_normal(g, i);
    surf[2 * static_cast<size_t>(i)] = make_float4(p.x, p.y, p.z, n.x);
    surf[2 * static_cast<size_t>(i) + 1] = make_float4(n.y, n.z, 0.0f, 0.0f);
}

template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) void k_pass_c_merge(const PassCArgs a, const CQueues q) {
    constexpr uint32_t kT = TX * TY;
    // neighbour pixel of slot j (j < k <= 16) of thread t at [j * kT + t]: read back by the same
    // wave only (its horizon tests and ray stores), which sees its own LDS writes in order
    __shared__ uint32_t nbr[16 * kT];
    uint32_t x, y;
    tile_pixel<TX, TY>(a.tilesX, a.nTiles, x, y);
    const uint32_t W = a.u.screenSize[0], H = a.u.screenSize[1];
    const bool inside = x < W && y < H;
    const bool vis = (a.u.flags & RESTIR_VISIBILITY_REUSE_FLAG) != 0;
    const size_t idx = inside ? static_cast<size_t>(y) * W + x : 0;
    // RESTIR_INPUT_VISIBLE_FLAG: a ray from the pixel whose input reservoir supplied the merged
    // sample retraces the segment that pixel's own centre re-test found unoccluded: not queued
    const bool known = (a.u.flags & RESTIR_INPUT_VISIBLE_FLAG) != 0;
    bool emit = false;    // the centre ray is queued
    uint32_t above = 0;   // above-horizon neighbour slots
    uint32_t queued = 0;  // the slots whose rays are queued
    f3 pos{0.0f, 0.0f, 0.0f}, lp{0.0f, 0.0f, 0.0f};
    if (inside) {
        const GView &g = a.g;
        pos = plane3(g.pos, idx, g.P);
        const f3 n = load_normal(g, idx);
        Res res = load_res(a.in, idx);
        if (dot(n, n) == 0.0f) {
            store_res(a.out, idx, res);
            q.bits[idx] = kBitBackground;
        } else {
            const uint32_t origM = res.M;
            uint32_t src = kNoPixel;
            pass_c_merge(
                a, x, y, idx, pos, n, res,
                [&](int i, uint32_t ni, uint32_t m) {
                    nbr[static_cast<uint32_t>(i) * kT + threadIdx.x] = ni;
                    q.nM[static_cast<size_t>(i) * q.npx + idx] = m;
                },
                &src);
            if (!known) src = kNoPixel;
            lp = res.pos;
            uint32_t numSamples = origM;
            // which neighbours see the merged sample above their horizon; unrolled for the usual
            // neighbour count, so that the surface loads of all slots are in flight together
            const auto horizon = [&](int j) {
                const uint32_t ni = nbr[static_cast<uint32_t>(j) * kT + threadIdx.x];
                const float4 s0 = q.surf[2 * static_cast<size_t>(ni)], s1 = q.surf[2 * static_cast<size_t>(ni) + 1];
                const f3 nwp{s0.x, s0.y, s0.z}, nn{s0.w, s1.x, s1.y};
                const uint32_t bit = dot(lp - nwp, nn) < 0.0f ? 0u : 1u << j;
                above |= bit;
                queued |= ni == src ? 0u : bit;
            };
            if (a.k == 5) {
#pragma unroll
                for (int j = 0; j < 5; ++j) horizon(j);
            } else {
                for (int j = 0; j < a.k; ++j) horizon(j);
            }
            if (!vis)
                for (uint32_t m = above; m; m &= m - 1)
                    numSamples += q.nM[static_cast<size_t>(__builtin_ctz(m)) * q.npx + idx];
            if (!vis)
                finish_w(res.w, res.sumW, res.pHat, numSamples, res.M,
                         (a.u.flags & RESTIR_STANDARD_WEIGHTS_FLAG) != 0);
            store_res(a.out, idx, res);
            q.origM[idx] = origM;
            q.occ[idx] = 0u;
            q.bits[idx] = above | (vis ? 0u : kBitDone);
            emit = vis && src != static_cast<uint32_t>(idx);
            if (!vis) queued = 0u;
        }
    }
    // The centre re-test (unbiasedReuse.glsl:157-166), and one neighbour ray per above-horizon
    // neighbour. The tracer skips the neighbour rays of a pixel whose centre ray it found shadowed:
    // a shadowed centre zeroes Z whatever the neighbours give (unbiasedReuse.glsl:157-166) and the
    // neighbour loop draws no random numbers. Both queues' room is reserved per workgroup. The
    // wave's neighbour rays are written slot-major: lane i writes rays i, i + 64, ... of the wave's
    // run, finding the pixel that owns a ray by binary search over the wave's prefix sums, so each
    // store instruction covers consecutive queue entries.
    const uint32_t want = queued;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t cnt = static_cast<uint32_t>(__popc(want));
    uint32_t incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= static_cast<uint32_t>(off)) incl += t;
    }
    const uint32_t excl = incl - cnt;
    const uint32_t total = __shfl(incl, 63);
    const uint2 rank = wave_rank(emit);
    __shared__ uint2 reserve[kT / 64 + 1];
    const uint2 first = group_reserve<kT / 64>(reinterpret_cast<unsigned long long *>(q.counter), rank.y, total, reserve);
    if (emit)
        q.c1[first.x + rank.x] =
            make_float4(pos.x, pos.y, pos.z, __uint_as_float(static_cast<uint32_t>(idx) | kCentreTag));
    if (total == 0) return;  // uniform: the whole wave leaves
    const uint32_t base = first.y;
    const uint32_t myIdx = static_cast<uint32_t>(idx);
    for (uint32_t r0 = 0; r0 < total; r0 += 64) {  // every lane runs every round
        const uint32_t r = r0 + lane;
        // owner: the last lane whose exclusive prefix is <= r (it has r < its inclusive prefix)
        uint32_t o = 0;
#pragma unroll
        for (uint32_t step = 32; step > 0; step >>= 1) {
            const uint32_t c = o + step;
            const uint32_t ex = __shfl(excl, static_cast<int>(c < 64 ? c : 63));
            if (c < 64 && ex <= r) o = c;
        }
        const uint32_t oWant = __shfl(want, static_cast<int>(o));
        const uint32_t oExcl = __shfl(excl, static_cast<int>(o));
        const uint32_t oIdx = __shfl(myIdx, static_cast<int>(o));
        if (r < total) {
            uint32_t m = oWant;
            for (uint32_t k = r - oExcl; k > 0; --k) m &= m - 1;  // the k-th set bit of the owner
            const uint32_t j = static_cast<uint32_t>(__ffs(static_cast<int>(m)) - 1);
            const float4 sp = q.surf[2 * static_cast<size_t>(nbr[j * kT + (threadIdx.x & ~63u) + o])];
            q.n1[base + r] = make_float4(sp.x, sp.y, sp.z, __uint_as_float(oIdx | (j << 27)));
        }
    }
}

template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) void k_pass_c_finish(const PassCArgs a, const CQueues q) {
    uint32_t x, y;
    tile_pixel<TX, TY>(a.tilesX, a.nTiles, x, y);
    const uint32_t W = a.u.screenSize[0], H = a.u.screenSize[1];
    if (x >= W || y >= H) return;
    const size_t idx = static_cast<size_t>(y) * W + x;
    const uint32_t b = q.bits[idx];
    if (b & (kBitBackground | kBitDone)) return;
    uint32_t numSamples = 0;
    if (!(b & kBitCentreShadowed)) {
        numSamples = q.origM[idx];
        for (uint32_t m = b & 0xffffu & ~q.occ[idx]; m; m &= m - 1) {
            const uint32_t j = static_cast<uint32_t>(__ffs(static_cast<int>(m)) - 1);
            numSamples += q.nM[static_cast<size_t>(j) * q.npx + idx];
        }
    }
    float4 *r = reinterpret_cast<float4 *>(a.out + idx) + 2;  // {lightIndex, pHat, sumW, w}
    float4 v = *r;
    const bool standard = (a.u.flags & RESTIR_STANDARD_WEIGHTS_FLAG) != 0;
    const uint32_t M = standard ? __float_as_uint(r[1].x) : 0u;  // numStreamSamples
    finish_w(v.w, v.z, v.y, numSamples, M, standard);
    *r = v;
}

// Tracer over a queue whose length a previous kernel left in *count. A fixed grid of workgroups
// (as many as fit the device at once); each wave takes chunks of kTraceChunk consecutive rays and
// hands a chunk's rays to its idle lanes in queue order whenever kTraceRefill or more lanes are
// idle. Queue order keeps the rays of neighbouring pixels together, which walk much of the same
// tree (the L1 serves most node fetches); the refills keep lanes busy when a wave's rays end at
// different depths. Chunk c belongs to shard c mod kTraceShards, whose counter (sc.traceHeads, one
// 128-byte line per shard) hands out its chunks in order, one atomic per chunk: a wave starts at
// its own shard and moves on to the next one when that is exhausted, so the atomics of the grid
// spread over all the lines (atomics on one line are served one at a time) while the chunks in
// flight stay near the front of the queue.
//
// Triangle tests are gathered per wave: a visit appends the triangles whose entries pass to the
// wave's LDS ring (triangle index and owner lane), and the wave tests the oldest 64, one per
// lane, whenever 64 are waiting, or all of them once fewer than kTriRoundWalkers lanes can still
// walk. A lane's ray ends when no test of its is waiting and a test hit or its walk is over; the
// lane takes no new ray before, and the tests of a ray already known to hit are skipped.
//
// MODE kTracePassA: the rays are Pass A's visibility rays (p1, p2: end points; *count of them),
// and one that is shadowed sets kBitCentreShadowed in bits[pixel]. kTracePassC: the rays are a
// Pass C's: its count[0] centre rays (q.c1), then its count[1] neighbour rays (q.n1), both
// ending at the light sample of merged[pixel]. A shadowed centre ray sets kBitCentreShadowed in
// bits[pixel]; a shadowed neighbour ray sets bit `slot` of occ[pixel], and is not traced when its
// pixel's centre ray is already known to be shadowed (the centre rays come first in the index
// space, so that holds for nearly all of them; tracing the others too changes nothing).
// kTraceBatch: nBatch rays p1[i] -> p2[i] (restir_trace_shadow); out[i] gets 1 if shadowed, else 0.
constexpr uint32_t kTraceRefill = 32;
constexpr uint32_t kTriList = kTraceListEntries, kTriRoundWalkers = 32;
constexpr int kTracePassA = 0, kTracePassC = 1, kTraceBatch = 2;
template <int STACK, int MODE>
__global__ __launch_bounds__(256) void k_trace_queue(const SceneView sc, const float4 *__restrict__ p1,
                                                     const float4 *__restrict__ p2,
                                                     const uint32_t *__restrict__ count, uint32_t nBatch, CQueues q,
                                                     uint32_t *__restrict__ out) {
    constexpr bool PER_PIXEL = MODE == kTracePassA, BATCH = MODE == kTraceBatch;
    // visits per loop step: two for Pass A's rays, where spreading the step's refill, round and
    // finish checks over two visits pays for the lower occupancy it costs (106 VGPRs, 4 waves
    // per SIMD); one for Pass C's, where it does not
    constexpr int kVisits = MODE == kTracePassA ? 2 : 1;
    __shared__ int32_t stack[STACK * 256];
    __shared__ uint32_t triList[4 * kTriList];
    __shared__ uint8_t triOwner[4 * kTriList];
    __shared__ uint32_t tally[256];  // per lane: 256 x (its tests done in a round) + its hits
    static_assert(sizeof(triList) + sizeof(triOwner) + sizeof(tally) == kTraceExtraLds, "LDS the grid is sized for");
    int32_t *stk = stack + threadIdx.x;
    const uint32_t nCentre = MODE == kTracePassC ? count[0] : 0u;
    const uint32_t n = BATCH ? nBatch : (PER_PIXEL ? count[0] : nCentre + count[1]);
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t *tri = triList + (threadIdx.x >> 6) * kTriList;
    uint8_t *own = triOwner + (threadIdx.x >> 6) * kTriList;
    uint32_t *tal = tally + (threadIdx.x & ~63u);
    tally[threadIdx.x] = 0u;  // each wave uses its own 64 words only
    WWalk w;
    int32_t ray = -1;
    uint32_t tag = 0;
    uint32_t waiting = 0;        // this lane's tests in the wave's list
    uint32_t first = 0, listed = 0;  // wave-uniform: the ring's oldest entry and its length
    uint32_t next = 0, end = 0;  // wave-uniform: the rays of the wave's chunk not handed out yet
    const uint32_t nChunks = (n + kTraceChunk - 1) / kTraceChunk;
    uint32_t shard = (blockIdx.x * 4u + (threadIdx.x >> 6)) % kTraceShards;  // wave-uniform
    uint32_t spent = 0;          // wave-uniform: shards this wave found exhausted
    bool drained = false;        // wave-uniform: every shard is exhausted
    for (;;) {
        const uint64_t idle = __ballot(ray < 0);
        const uint32_t nIdle = static_cast<uint32_t>(__popcll(idle));
        if (nIdle >= kTraceRefill) {
            if (next >= end && !drained) {
                // every lane is active here: the first active lane is lane 0, and the values stay
                // in scalar registers, so the loop's tests are seen as wave-uniform
                for (;;) {
                    uint32_t *h = sc.traceHeads + shard * kTraceShardStride;
                    // a shard already past its last chunk is passed over without an atomic
                    uint32_t t = __builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                    if (t * kTraceShards + shard < nChunks) {
                        uint32_t v = 0;
                        if (lane == 0) v = atomicAdd(h, 1u);
                        t = __builtin_amdgcn_readfirstlane(v);
                    }
                    const uint32_t c = t * kTraceShards + shard;
                    if (c < nChunks) {
                        next = c * kTraceChunk;
                        end = n - next < kTraceChunk ? n : next + kTraceChunk;
                        break;
                    }
                    if (++spent == kTraceShards) {
                        drained = true;
                        break;
                    }
                    shard = shard + 1 == kTraceShards ? 0u : shard + 1;
                }
            }
            if (next < end) {
                if (ray < 0) {
                    const uint32_t i = next + lanes_below(idle);
                    if (i < end) {
                        const bool centre = MODE == kTracePassC && i < nCentre;
                        const float4 a4 = MODE != kTracePassC ? p1[i] : (centre ? q.c1[i] : q.n1[i - nCentre]);
                        tag = BATCH ? i : __float_as_uint(a4.w);
                        // a neighbour ray of a pixel whose centre ray was shadowed is not traced
                        if (MODE != kTracePassC || centre || !(q.bits[tag & 0x07ffffffu] & kBitCentreShadowed)) {
                            f3 o, d;
                            const float4 b4 = MODE != kTracePassC
                                                  ? p2[i]
                                                  : reinterpret_cast<const float4 *>(q.merged + (tag & 0x07ffffffu))[0];
                            shadow_segment(xyz(a4), xyz(b4), o, d);
                            wwalk_begin(w, o, d, sc.treeNested != 0, sc.maxCoord);
                            ray = static_cast<int32_t>(i);
                        }
                    }
                }
                next = end - next < nIdle ? end : next + nIdle;
            }
        }
        if (__ballot(ray >= 0) == 0) {
            if (drained && next >= end) break;
            continue;
        }
        // kVisits visits per walking lane (the second only while the list holds fewer than 64
        // tests); their triangles join the wave's list
#pragma unroll
        for (int rep = 0; rep < kVisits; ++rep) {
            if (rep > 0 && listed >= 64u) break;  // wave-uniform
            const bool walking = ray >= 0 && !w.hit && w.cur >= 0;
            uint32_t leaves = 0;
            int4 ids;  // read only where leaves has bits, that is after a visit wrote it
            if (walking) leaves = walk_step_defer<STACK>(w, sc.wide, sc.nodes, stk, 256, ids);
            const uint32_t nLeaves = static_cast<uint32_t>(__popc(leaves));
            const uint64_t c1 = __ballot(nLeaves & 1u), c2 = __ballot(nLeaves & 2u), c4 = __ballot(nLeaves & 4u);
            if (c1 | c2 | c4) {
                // listed < 64 before the visit, so the ring (kTriList entries) cannot overflow
                uint32_t at = first + listed + lanes_below(c1) + 2u * lanes_below(c2) + 4u * lanes_below(c4);
                // one step per entry, so that each entry's id is a register rather than a select
                const auto put = [&](int32_t id) {
                    const uint32_t slot = at < kTriList ? at : at - kTriList;
                    tri[slot] = static_cast<uint32_t>(~id);
                    own[slot] = static_cast<uint8_t>(lane);
                    ++at;
                };
                if (leaves & 1u) put(ids.x);
                if (leaves & 2u) put(ids.y);
                if (leaves & 4u) put(ids.z);
                if (leaves & 8u) put(ids.w);
                listed += static_cast<uint32_t>(__popcll(c1) + 2 * __popcll(c2) + 4 * __popcll(c4));
                waiting += nLeaves;
            }
        }
        // triangle rounds: the oldest 64 tests, one per lane
        const uint32_t walkers = static_cast<uint32_t>(__popcll(__ballot(ray >= 0 && !w.hit && w.cur >= 0)));
        while (listed >= 64u || (listed > 0u && walkers < kTriRoundWalkers)) {
            const uint64_t known = __ballot(ray >= 0 && w.hit);
            const uint32_t take = listed < 64u ? listed : 64u;
            const bool mine = lane < take;
            uint32_t t = 0, o = 0;
            if (mine) {
                const uint32_t at = first + lane, slot = at < kTriList ? at : at - kTriList;
                t = tri[slot];
                o = own[slot];
            }
            const int src = static_cast<int>(o);
            const f3 ro{__shfl(w.s.s.o.x, src), __shfl(w.s.s.o.y, src), __shfl(w.s.s.o.z, src)};
            const f3 rd{__shfl(w.s.s.d.x, src), __shfl(w.s.s.d.y, src), __shfl(w.s.s.d.z, src)};
            bool h = false;
            if (mine && !((known >> o) & 1ull)) h = ray_tri(sc.tris, static_cast<int32_t>(t), ro, rd);
            if (mine) atomicAdd(&tal[o], h ? 257u : 256u);
            const uint32_t r = atomicExch(&tal[lane], 0u);  // after this wave's adds (LDS keeps a wave's order)
            waiting -= r >> 8;
            if (r & 0xffu) w.hit = true;
            first = first + take < kTriList ? first + take : first + take - kTriList;
            listed -= take;
        }
        if (ray >= 0 && waiting == 0 && (w.hit || w.cur < 0)) {
            if (BATCH) {
                out[tag] = w.hit ? 1u : 0u;
            } else if (w.hit) {
                if (PER_PIXEL || (tag & kCentreTag)) q.bits[tag & 0x07ffffffu] |= kBitCentreShadowed;
                else atomicOr(&q.occ[tag & 0x07ffffffu], 1u << (tag >> 27));
            }
            ray = -1;
        }
    }
}

// ---------------------------------------------------------------- Pass A, ray queue ---
// Pass A with its visibility ray (restirOmni.glsl:148-160) moved out of the per-pixel lanes:
// (1) RIS per pixel; the reservoir is parked in the output, the random stream's state in
// scratch, and the ray queued; (2) the persistent tracer marks shadowed pixels; (3) per pixel:
// the shadow verdict, temporal reuse, store. Same draws, same order, same result as k_pass_a.
struct AQueues {
    uint32_t *counter;  // [0] rays queued
    float4 *p1, *p2;    // p1.w carries the pixel index
    uint64_t *rng;      // per pixel: the random stream's state after RIS
    uint32_t *occ;      // per pixel: kBitCentreShadowed once its ray hit
};

template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) void k_pass_a_ris(const PassAArgs a, const AQueues q) {
    uint32_t x, y;
    tile_pixel<TX, TY>(a.tilesX, a.nTiles, x, y);
    const bool inside = x < a.u.screenSize[0] && y < a.u.screenSize[1];
    bool emit = false;
    size_t idx = 0;
    f3 pos{0.0f, 0.0f, 0.0f}, lp{0.0f, 0.0f, 0.0f};
    if (inside) {
        const PassASurface ps = pass_a_surface(a, x, y);
        idx = ps.idx;
        pos = ps.pos;
        Res res;
        Pcg rng = pass_a_rng(a, x, y);
        pass_a_ris(a, ps, res, rng);
        store_res(a.out, idx, res);
        q.rng[idx] = rng.state;
        q.occ[idx] = 0u;
        emit = (a.u.flags & RESTIR_VISIBILITY_REUSE_FLAG) && res.sumW != 0.0f;
        lp = res.pos;
    }
    __shared__ uint2 reserve[TX * TY / 64 + 1];
    const uint2 rank = wave_rank(emit);
    const uint32_t slot = group_reserve<TX * TY / 64>(reinterpret_cast<unsigned long long *>(q.counter), rank.y, 0u,
                                                      reserve).x + rank.x;
    if (emit) {
        q.p1[slot] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(static_cast<uint32_t>(idx)));
        q.p2[slot] = make_float4(lp.x, lp.y, lp.z, 0.0f);
    }
}

template <int TX, int TY>
__global__ __launch_bounds__(TX *TY) void k_pass_a_finish(const PassAArgs a, const AQueues q) {
    uint32_t x, y;
    tile_pixel<TX, TY>(a.tilesX, a.nTiles, x, y);
    if (x >= a.u.screenSize[0] || y >= a.u.screenSize[1]) return;
    const PassASurface ps = pass_a_surface(a, x, y);
    Res res = load_res(a.out, ps.idx);
    Pcg rng = pass_a_rng(a, x, y);  // the stream's increment; its state resumes after RIS
    rng.state = q.rng[ps.idx];
    if ((a.u.flags & RESTIR_VISIBILITY_REUSE_FLAG) && res.sumW != 0.0f && (q.occ[ps.idx] & kBitCentreShadowed)) {
        res.w = 0.0f;
        res.sumW = 0.0f;
    }
    pass_a_temporal(a, ps, res, rng);
    store_res(a.out, ps.idx, res);
}

// ------------------------------------------------------------------------------- Shading --
struct ShadeArgs {
    SceneView sc;
    f3 cam;
    float gamma;
    GView g;
    const restir_reservoir *res;
    float *rgb;
    uint64_t rgbP;
};

// occluded: the pixel's sample failed the final visibility re-test (it then shades to 0)
__device__ __forceinline__ void shade_px(const ShadeArgs &a, size_t i, bool occluded = false) {
    const GView &g = a.g;
    if (i >= static_cast<size_t>(g.W) * g.H) return;
    const f3 pos = plane3(g.pos, i, g.P);
    const f3 n = load_normal(g, i);
    const f3 alb = plane3(g.alb, i, g.P);
    const float emissive = g.alb[3 * g.P + i];
    f3 c{0.0f, 0.0f, 0.0f};
    if (dot(n, n) != 0.0f) {
        const Res r = load_res(a.res, i);
        // a light index outside the tables (a foreign buffer) shades black instead of faulting
        const float kZero[3] = {0.0f, 0.0f, 0.0f};
        const float *em = kZero;
        if (r.lightIndex < 0) {
            if (-1 - r.lightIndex < a.sc.nTriLights) em = a.sc.triLights[-1 - r.lightIndex].emission_luminance;
        } else if (r.lightIndex < a.sc.nPoints) {
            em = a.sc.points[r.lightIndex].color_luminance;
        }
        const Surface s = make_surface(pos, n, a.cam, 0.0f, g.mat[i], g.mat[g.P + i]);
        const f3 f = eval_phat_full(s, alb, r.pos, f3{r.normal.x, r.normal.y, r.normal.z},
                                    r.normal.w > 0.5f, f3{em[0], em[1], em[2]});
        c = f3{f.x * r.w, f.y * r.w, f.z * r.w};
        if (occluded) c = f3{0.0f, 0.0f, 0.0f};
    }
    if (emissive > 0.5f) c = alb;
    if (a.gamma != 1.0f) {
        const float e = 1.0f / a.gamma;
        c = f3{powf(c.x, e), powf(c.y, e), powf(c.z, e)};
    }
    a.rgb[i] = c.x;
    a.rgb[a.rgbP + i] = c.y;
    a.rgb[2 * a.rgbP + i] = c.z;
}

// VIS: the visibility re-test of the final shading (the 1080p workload's setting): a pixel whose
// sample has nonzero W traces the segment from its surface to the sample (testVisibility) first,
// one ray per lane.
template <int STACK, bool VIS>
__global__ __launch_bounds__(256) void k_shade(const ShadeArgs a) {
    __shared__ int32_t stack[VIS ? STACK * 256 : 1];
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    bool occ = false;
    if (VIS && i < static_cast<size_t>(a.g.W) * a.g.H) {
        const f3 n = load_normal(a.g, i);
        if (dot(n, n) != 0.0f) {
            const Res r = load_res(a.res, i);
            if (r.w != 0.0f) {
                f3 o, d;
                shadow_segment(plane3(a.g.pos, i, a.g.P), r.pos, o, d);
                TraceStats st{0, 0};
                occ = occluded_w<STACK>(a.sc.wide, a.sc.nodes, a.sc.tris, a.sc.treeNested != 0, a.sc.maxCoord, o, d,
                                        stack + threadIdx.x, 256, st);
            }
        }
    }
    shade_px(a, i, occ);
}

// ------------------------------------------------------------------------- batch trace ----
struct TraceArgs {
    SceneView sc;
    const float4 *p1, *p2;
    uint32_t n;
    uint32_t *shadowed, *visits, *tests;
};

template <int STACK, bool REF>
__global__ __launch_bounds__(256) void k_trace(const TraceArgs a) {
    __shared__ int32_t stack[STACK * 256];
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    f3 o, d;
    shadow_segment(xyz(a.p1[i]), xyz(a.p2[i]), o, d);
    TraceStats st{0, 0};
    bool occ;
    if constexpr (REF) occ = occluded<STACK, true>(a.sc.nodes, a.sc.tris, o, d, stack + threadIdx.x, 256, st);
    else occ = occluded_w<STACK>(a.sc.wide, a.sc.nodes, a.sc.tris, a.sc.treeNested != 0, a.sc.maxCoord, o, d, stack + threadIdx.x,
                                 256, st);
    a.shadowed[i] = occ ? 1u : 0u;
    if (a.visits) a.visits[i] = st.visits;
    if (a.tests) a.tests[i] = st.tests;
}

// Persistent schedule: a fixed grid of workgroups pulls rays from a global counter. Each lane
// walks one node per step; when half a wave (or all of it) is idle, the wave claims that many
// rays with a single atomic and hands them out by prefix rank over the ballot of idle lanes, so
// lanes whose rays ended early do not wait for the slowest ray of their wave.
template <int STACK, bool REF>
__global__ __launch_bounds__(256) void k_trace_pool(const TraceArgs a, uint32_t *__restrict__ head) {
    __shared__ int32_t stack[STACK * 256];
    int32_t *stk = stack + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    std::conditional_t<REF, Walk, WWalk> w;  // the reference loop walks the binary nodes
    int32_t ray = -1;
    bool drained = false;  // wave-uniform: the counter has passed the last ray
    for (;;) {
        const uint64_t idle = __ballot(ray < 0);
        const uint32_t nIdle = static_cast<uint32_t>(__popcll(idle));
        if (!drained && nIdle >= 32u) {
            const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<long long>(idle)) - 1);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(head, nIdle);
            base = __shfl(base, static_cast<int>(leader));
            drained = base + nIdle >= a.n;
            if (ray < 0) {
                const uint32_t i = base + lanes_below(idle);
                if (i < a.n) {
                    f3 o, d;
                    shadow_segment(xyz(a.p1[i]), xyz(a.p2[i]), o, d);
                    if constexpr (REF) walk_begin(w, o, d);
                    else wwalk_begin(w, o, d, a.sc.treeNested != 0, a.sc.maxCoord);
                    ray = static_cast<int32_t>(i);
                }
            }
        }
        if (__ballot(ray >= 0) == 0) {
            if (drained) break;
            continue;
        }
        bool fin = false;
        if (ray >= 0) {
            if constexpr (REF) fin = walk_step<STACK, true>(w, a.sc.nodes, a.sc.tris, stk, 256);
            else fin = walk_step_w<STACK>(w, a.sc.wide, a.sc.nodes, a.sc.tris, stk, 256);
        }
        if (fin) {
            a.shadowed[ray] = w.hit ? 1u : 0u;
            if (a.visits) a.visits[ray] = w.nv;
            if (a.tests) a.tests[ray] = w.nt;
            ray = -1;
        }
    }
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ void k_zero_counters(uint32_t *p, uint32_t n, uint32_t *heads) {
    if (threadIdx.x < n) p[threadIdx.x] = 0u;
    if (threadIdx.x < kTraceShards) heads[threadIdx.x * kTraceShardStride] = 0u;
}

// ---------------------------------------------------------------------- primary G-buffer --
struct PrimaryArgs {
    SceneView sc;
    CamBasis cb;
    const int32_t *triMaterial;
    const float *materials;
    float *pos;
    void *nrm;  // float planes, or half planes when nrmHalf
    float *alb, *mat, *depth;
    uint64_t P;
    uint32_t W, H;
    bool nrmHalf;
};

// Writes normal n of pixel i as floats or as halves (round to nearest even).
__device__ __forceinline__ void store_normal(const PrimaryArgs &a, size_t i, f3 n) {
    const uint64_t P = a.P;
    if (a.nrmHalf) {
        _Float16 *h = static_cast<_Float16 *>(a.nrm);
        h[i] = static_cast<_Float16>(n.x);
        h[P + i] = static_cast<_Float16>(n.y);
        h[2 * P + i] = static_cast<_Float16>(n.z);
    } else {
        float *f = static_cast<float *>(a.nrm);
        f[i] = n.x;
        f[P + i] = n.y;
        f[2 * P + i] = n.z;
    }
}

// Closest-hit slab test on a ray (not a segment): entry distance, or NaN on a miss (so every
// comparison against it is false).
__device__ __forceinline__ float box_entry(f3 o, f3 d, float4 mn, float4 mx) {
    const float ax = (mn.x - o.x) / d.x, ay = (mn.y - o.y) / d.y, az = (mn.z - o.z) / d.z;
    const float bx = (mx.x - o.x) / d.x, by = (mx.y - o.y) / d.y, bz = (mx.z - o.z) / d.z;
    const float rmin = gmax(gmin(ax, bx), gmax(gmin(ay, by), gmin(az, bz)));
    const float rmax = gmin(gmax(ax, bx), gmin(gmax(ay, by), gmax(az, bz)));
    return (rmax >= rmin && rmax > 0.0f) ? rmin : __builtin_nanf("");
}

__device__ __forceinline__ bool tri_closest(const float4 *__restrict__ tris, int32_t t, f3 o, f3 d,
                                            float &tout) {
    const float4 a = tris[3 * t + 0], b = tris[3 * t + 1], c = tris[3 * t + 2];
    const f3 p1 = xyz(a);
    const f3 e1 = xyz(b) - p1;
    const f3 e2 = xyz(c) - p1;
    const f3 p = cross(d, e2);
    const float f = 1.0f / dot(e1, p);
    const f3 s = o - p1;
    const float bx = f * dot(s, p);
    if (bx < 0.0f || bx > 1.0f) return false;
    const f3 q = cross(s, e1);
    const float by = f * dot(d, q);
    if (by < 0.0f || by + bx > 1.0f) return false;
    const float tt = f * dot(e2, q);
    if (!(tt > 0.0f)) return false;
    tout = tt;
    return true;
}

template <int STACK>
__global__ __launch_bounds__(256) void k_primary(const PrimaryArgs a) {
    __shared__ int32_t stack[STACK * 256];
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= static_cast<size_t>(a.W) * a.H) return;
    const uint32_t x = static_cast<uint32_t>(i % a.W), y = static_cast<uint32_t>(i / a.W);
    const CamBasis &b = a.cb;
    const float ndx = (static_cast<float>(x) + 0.5f) / static_cast<float>(a.W) * 2.0f - 1.0f;
    const float ndy = (static_cast<float>(y) + 0.5f) / static_cast<float>(a.H) * 2.0f - 1.0f;
    const f3 d = normalize((b.fwd + b.right * (ndx * b.sx)) + b.down * (ndy * b.sy));
    const f3 o = b.pos;
    float best = __builtin_huge_valf();
    int32_t bi = -1;
    int32_t *stk = stack + threadIdx.x;
    int top = 0;
    int32_t cur = 0;
    for (;;) {
        if (cur < 0) {
            if (top == 0) break;
            --top;
            cur = stk[top * 256];
        }
        const Node nd = load_node(a.sc.nodes, cur);
        // conservative culling: entry distances are compared with a relative slack so rounding
        // of the slab test never discards a triangle whose computed t ties the best hit
        const float lim = best * 1.00001f + 1e-6f;
        const float tl = box_entry(o, d, nd.lmin, nd.lmax);
        const float tr = box_entry(o, d, nd.rmin, nd.rmax);
        const bool li = tl <= lim, ri = tr <= lim;
        int32_t inner[2];
        int ni = 0;
        if (li) {
            if (nd.left < 0) {
                float t;
                const int32_t ti = ~nd.left;
                if (tri_closest(a.sc.tris, ti, o, d, t) && (t < best || (t == best && ti < bi))) {
                    best = t;
                    bi = ti;
                }
            } else {
                inner[ni++] = nd.left;
            }
        }
        if (ri) {
            if (nd.right < 0) {
                float t;
                const int32_t ti = ~nd.right;
                if (tri_closest(a.sc.tris, ti, o, d, t) && (t < best || (t == best && ti < bi))) {
                    best = t;
                    bi = ti;
                }
            } else {
                inner[ni++] = nd.right;
            }
        }
        if (ni == 2) {
            // visit the nearer box first
            const bool leftFirst = tl <= tr;
            const int32_t first = leftFirst ? inner[0] : inner[1];
            const int32_t second = leftFirst ? inner[1] : inner[0];
            if (top < STACK) {
                stk[top * 256] = second;
                ++top;
            }
            cur = first;
        } else if (ni == 1) {
            cur = inner[0];
        } else {
            cur = -1;
        }
    }
    const uint64_t P = a.P;
    if (bi < 0) {
        a.pos[i] = a.pos[P + i] = a.pos[2 * P + i] = 0.0f;
        store_normal(a, i, f3{0.0f, 0.0f, 0.0f});
        a.alb[i] = a.alb[P + i] = a.alb[2 * P + i] = a.alb[3 * P + i] = 0.0f;
        a.mat[i] = a.mat[P + i] = 0.0f;
        a.depth[i] = 1.0f;
        return;
    }
    const f3 hp = o + d * best;
    const float4 ta = a.sc.tris[3 * bi], tb = a.sc.tris[3 * bi + 1], tc = a.sc.tris[3 * bi + 2];
    const f3 p1 = xyz(ta);
    f3 nn = normalize(cross(xyz(tb) - p1, xyz(tc) - p1));
    if (dot(nn, d) > 0.0f) nn = neg(nn);
    const float *m = a.materials + 8 * static_cast<size_t>(a.triMaterial[bi]);
    a.pos[i] = hp.x;
    a.pos[P + i] = hp.y;
    a.pos[2 * P + i] = hp.z;
    store_normal(a, i, nn);
    a.alb[i] = m[0];
    a.alb[P + i] = m[1];
    a.alb[2 * P + i] = m[2];
    a.alb[3 * P + i] = m[3];
    a.mat[i] = m[4];
    a.mat[P + i] = m[5];
    const float z = best * dot(d, b.fwd);
    a.depth[i] = (b.A * z + b.B) / z;
}

// ----------------------------------------------------------------------------- launchers --
template <int TX, int TY>
static void grid_for(uint32_t W, uint32_t H, uint32_t &tilesX, uint32_t &nTiles) {
    tilesX = (W + TX - 1) / TX;
    nTiles = tilesX * ((H + TY - 1) / TY);
}

template <int TX, int TY, int STACK>
static hipError_t run_a(PassAArgs args, hipStream_t s) {
    grid_for<TX, TY>(args.u.screenSize[0], args.u.screenSize[1], args.tilesX, args.nTiles);
    RK_LAUNCH("k_pass_a", (k_pass_a<TX, TY, STACK>), dim3(args.nTiles), dim3(TX * TY), 0, s, args);
    return hipGetLastError();
}
template <int STACK>
static hipError_t dispatch_a(int tile, const PassAArgs &args, hipStream_t s) {
    switch (tile) {
        case 0: return run_a<16, 16, STACK>(args, s);
        case 1: return run_a<64, 1, STACK>(args, s);
        case 2: return run_a<8, 8, STACK>(args, s);
        case 3: return run_a<16, 8, STACK>(args, s);
        default: return run_a<32, 8, STACK>(args, s);
    }
}
static AQueues carve_a_queues(void *scratch, uint32_t npx, size_t *bytes = nullptr) {
    Carver c(scratch);
    const size_t n = npx;
    AQueues q;
    q.counter = c.take<uint32_t>(256);
    q.p1 = c.take<float4>(sizeof(float4) * n);
    q.p2 = c.take<float4>(sizeof(float4) * n);
    q.rng = c.take<uint64_t>(sizeof(uint64_t) * n);
    q.occ = c.take<uint32_t>(sizeof(uint32_t) * n);
    if (bytes) *bytes = c.end();
    return q;
}
size_t pass_a_queue_bytes(uint32_t npx) {
    size_t bytes = 0;
    carve_a_queues(nullptr, npx, &bytes);
    return bytes;
}

template <int TX, int TY, int STACK>
static hipError_t run_a_queued(PassAArgs args, const AQueues &q, uint32_t gridCap, hipStream_t s) {
    grid_for<TX, TY>(args.u.screenSize[0], args.u.screenSize[1], args.tilesX, args.nTiles);
    const dim3 tiles(args.nTiles), block(TX * TY);
    RK_LAUNCH("k_zero_counters", k_zero_counters, dim3(1), dim3(64), 0, s, q.counter, 2u, args.sc.traceHeads);
    RK_LAUNCH("k_pass_a_ris", (k_pass_a_ris<TX, TY>), tiles, block, 0, s, args, q);
    if (args.u.flags & RESTIR_VISIBILITY_REUSE_FLAG) {
        CQueues cq{};
        cq.bits = q.occ;  // a shadowed ray sets kBitCentreShadowed in occ[pixel]
        RK_LAUNCH("k_trace_queue(pass_a)", (k_trace_queue<STACK, kTracePassA>), dim3(gridCap), dim3(256), 0, s,
                  args.sc, q.p1, q.p2, q.counter + 0, 0u, cq, nullptr);
    }
    RK_LAUNCH("k_pass_a_finish", (k_pass_a_finish<TX, TY>), tiles, block, 0, s, args, q);
    return hipGetLastError();
}
template <int STACK>
static hipError_t dispatch_a_queued(int tile, const PassAArgs &args, const AQueues &q, uint32_t gridCap,
                                    hipStream_t s) {
    switch (tile) {
        case 0: return run_a_queued<16, 16, STACK>(args, q, gridCap, s);
        case 1: return run_a_queued<64, 1, STACK>(args, q, gridCap, s);
        case 2: return run_a_queued<8, 8, STACK>(args, q, gridCap, s);
        case 3: return run_a_queued<16, 8, STACK>(args, q, gridCap, s);
        default: return run_a_queued<32, 8, STACK>(args, q, gridCap, s);
    }
}

hipError_t launch_pass_a(int tile, int stack, const restir_uniforms &u, const SceneView &sc,
                         const GView &cur, const GView &prev, const restir_reservoir *prevRes,
                         restir_reservoir *out, const int32_t *candLight, const float *candQ,
                         void *queueScratch, uint32_t gridCap, hipStream_t s) {
    PassAArgs args{u, sc, cur, prev, prevRes, out, 0, 0, candLight, candQ};
    if (queueScratch) {
        const AQueues q = carve_a_queues(queueScratch, u.screenSize[0] * u.screenSize[1]);
        return by_stack(stack, [&](auto S) {
            return dispatch_a_queued<decltype(S)::value>(tile, args, q, gridCap, s);
        });
    }
    return by_stack(stack, [&](auto S) { return dispatch_a<decltype(S)::value>(tile, args, s); });
}

template <int TX, int TY>
static hipError_t run_b(PassBArgs args, hipStream_t s) {
    grid_for<TX, TY>(args.u.screenSize[0], args.u.screenSize[1], args.tilesX, args.nTiles);
    RK_LAUNCH("k_pass_b", (k_pass_b<TX, TY>), dim3(args.nTiles), dim3(TX * TY), 0, s, args);
    return hipGetLastError();
}
hipError_t launch_pass_b(int tile, const restir_uniforms &u, int iter, float cosThr,
                         const GView &g, const restir_reservoir *in, restir_reservoir *out,
                         hipStream_t s) {
    PassBArgs args{u, g, in, out, 0, 0, u.frame * 31u + static_cast<uint32_t>(iter), cosThr};
    switch (tile) {
        case 0: return run_b<16, 16>(args, s);
        case 1: return run_b<64, 1>(args, s);
        case 2: return run_b<8, 8>(args, s);
        case 3: return run_b<16, 8>(args, s);
        default: return run_b<32, 8>(args, s);
    }
}

template <int TX, int TY, int STACK>
static hipError_t run_c(PassCArgs args, hipStream_t s) {
    grid_for<TX, TY>(args.u.screenSize[0], args.u.screenSize[1], args.tilesX, args.nTiles);
    RK_LAUNCH("k_pass_c", (k_pass_c<TX, TY, STACK>), dim3(args.nTiles), dim3(TX * TY), 0, s, args);
    return hipGetLastError();
}
template <int STACK>
static hipError_t dispatch_c(int tile, const PassCArgs &args, hipStream_t s) {
    switch (tile) {
        case 0: return run_c<16, 16, STACK>(args, s);
        case 1: return run_c<64, 1, STACK>(args, s);
        case 2: return run_c<8, 8, STACK>(args, s);
        case 3: return run_c<16, 8, STACK>(args, s);
        default: return run_c<32, 8, STACK>(args, s);
    }
}
template <int TX, int TY, int STACK>
static hipError_t run_c_queued(PassCArgs args, const CQueues &q, uint32_t gridCap, bool packSurface,
                               hipStream_t s) {
    static_assert((kCentreTag & 0x07ffffffu) == 0 && (kCentreTag >> 27) > 15u, "the centre tag is no pixel or slot");
    grid_for<TX, TY>(args.u.screenSize[0], args.u.screenSize[1], args.tilesX, args.nTiles);
    const dim3 tiles(args.nTiles), block(TX * TY);
    RK_LAUNCH("k_zero_counters", k_zero_counters, dim3(1), dim3(64), 0, s, q.counter, 2u, args.sc.traceHeads);
    const uint32_t npx = args.u.screenSize[0] * args.u.screenSize[1];
    // packSurface false: q.surf already holds this G-buffer's records (the previous Pass C's)
    if (packSurface)
        RK_LAUNCH("k_pack_surface", k_pack_surface, dim3((npx + 255) / 256), dim3(256), 0, s, args.g, npx, q.surf);
    RK_LAUNCH("k_pass_c_merge", (k_pass_c_merge<TX, TY>), tiles, block, 0, s, args, q);
    if (args.u.flags & RESTIR_VISIBILITY_REUSE_FLAG) {
        RK_LAUNCH("k_trace_queue(pass_c)", (k_trace_queue<STACK, kTracePassC>), dim3(gridCap), dim3(256), 0, s,
                  args.sc, nullptr, nullptr, q.counter + 0, 0u, q, nullptr);
        RK_LAUNCH("k_pass_c_finish", (k_pass_c_finish<TX, TY>), tiles, block, 0, s, args, q);
    }
    return hipGetLastError();
}
template <int STACK>
static hipError_t dispatch_c_queued(int tile, const PassCArgs &args, const CQueues &q,
                                    uint32_t gridCap, bool packSurface, hipStream_t s) {
    switch (tile) {
        case 0: return run_c_queued<16, 16, STACK>(args, q, gridCap, packSurface, s);
        case 1: return run_c_queued<64, 1, STACK>(args, q, gridCap, packSurface, s);
        case 2: return run_c_queued<8, 8, STACK>(args, q, gridCap, packSurface, s);
        case 3: return run_c_queued<16, 8, STACK>(args, q, gridCap, packSurface, s);
        default: return run_c_queued<32, 8, STACK>(args, q, gridCap, packSurface, s);
    }
}

// Carves the queues out of one scratch allocation of pass_c_queue_bytes(npx, k) bytes.
static CQueues carve_queues(void *scratch, uint32_t npx, int k, size_t *bytes = nullptr) {
    Carver c(scratch);
    const size_t n = npx, kk = static_cast<size_t>(k);
    CQueues q{};
    q.counter = c.take<uint32_t>(256);
    q.c1 = c.take<float4>(sizeof(float4) * n);
    q.n1 = c.take<float4>(sizeof(float4) * n * kk);
    q.origM = c.take<uint32_t>(sizeof(uint32_t) * n);
    q.bits = c.take<uint32_t>(sizeof(uint32_t) * n);
    q.occ = c.take<uint32_t>(sizeof(uint32_t) * n);
    q.nM = c.take<uint32_t>(sizeof(uint32_t) * n * kk);
    q.surf = c.take<float4>(2 * sizeof(float4) * n);
    q.npx = npx;
    if (bytes) *bytes = c.end();
    return q;
}
size_t pass_c_queue_bytes(uint32_t npx, int k) {
    size_t bytes = 0;
    carve_queues(nullptr, npx, k, &bytes);
    return bytes;
}

__global__ void k_traced_flag(const CQueues q, uint32_t len, uint32_t *flag) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= len) return;
    const uint32_t tag = __float_as_uint(q.n1[i].w);
    flag[i] = (q.bits[tag & 0x07ffffffu] & kBitCentreShadowed) ? 0u : 1u;
}
__global__ void k_centre_copy(const CQueues q, uint32_t len, float4 *p1, float4 *p2) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= len) return;
    const float4 a = q.c1[i];
    const uint32_t px = __float_as_uint(a.w) & 0x07ffffffu;
    const float4 b = reinterpret_cast<const float4 *>(q.merged + px)[0];
    p1[i] = make_float4(a.x, a.y, a.z, __uint_as_float(px));
    p2[i] = make_float4(b.x, b.y, b.z, 0.0f);
}
__global__ void k_traced_copy(const CQueues q, uint32_t len, const uint32_t *flag, const uint32_t *off, float4 *p1,
                              float4 *p2, uint32_t cap) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= len || !flag[i] || off[i] >= cap) return;
    const float4 a = q.n1[i];
    const float4 b = reinterpret_cast<const float4 *>(q.merged + (__float_as_uint(a.w) & 0x07ffffffu))[0];
    p1[off[i]] = a;
    p2[off[i]] = make_float4(b.x, b.y, b.z, 0.0f);
}

hipError_t pass_c_rays(const void *queueScratch, uint32_t npx, int k, const restir_reservoir *merged, int which,
                       float4 *p1, float4 *p2, uint32_t cap, uint32_t *n, hipStream_t s) {
    CQueues q = carve_queues(const_cast<void *>(queueScratch), npx, k);
    q.merged = merged;
    uint32_t len = 0;
    hipError_t e = hipMemcpyAsync(&len, q.counter + which, sizeof len, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const uint32_t maxLen = which == 0 ? npx : npx * static_cast<uint32_t>(k);
    if (len > maxLen) return hipErrorIllegalState;
    *n = len;
    if (which == 0) {
        *n = len;
        const uint32_t m = len < cap ? len : cap;
        if (m == 0) return hipSuccess;
        RK_LAUNCH("k_centre_copy", k_centre_copy, dim3((m + 255) / 256), dim3(256), 0, s, q, m, p1, p2);
        return hipGetLastError();
    }
    // neighbour rays: the ones the tracer traced (centre ray not shadowed), in queue order
    *n = 0;
    if (len == 0) return hipSuccess;
    uint32_t *tmp = nullptr;
    const size_t words = 2 * static_cast<size_t>(len) + rscan::temp_items(len) + 1;
    e = hipMalloc(&tmp, sizeof(uint32_t) * words);
    if (e != hipSuccess) return e;
    uint32_t *flag = tmp, *off = tmp + len, *temp = tmp + 2 * static_cast<size_t>(len), *total = tmp + words - 1;
    const uint32_t blocks = (len + 255) / 256;
    RK_LAUNCH("k_traced_flag", k_traced_flag, dim3(blocks), dim3(256), 0, s, q, len, flag);
    e = rscan::exclusive_scan<uint32_t>(flag, off, len, temp, total, s);
    if (e == hipSuccess) {
        RK_LAUNCH("k_traced_copy", k_traced_copy, dim3(blocks), dim3(256), 0, s, q, len, flag, off, p1, p2, cap);
        e = hipGetLastError();
    }
    uint32_t kept = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&kept, total, sizeof kept, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    const hipError_t ef = hipFree(tmp);
    if (e == hipSuccess) e = ef;
    if (e == hipSuccess) *n = kept;
    return e;
}

hipError_t launch_pass_c(int tile, int stack, const restir_uniforms &u, int k, int iter,
                         const SceneView &sc, const GView &g, const restir_reservoir *in,
                         restir_reservoir *out, void *queueScratch, uint32_t gridCap, hipStream_t s,
                         bool packSurface) {
    // unbiasedReuse.glsl:72 seeds with frame*17; iter > 0 offsets the seed of repeated passes
    PassCArgs args{u, sc, g, in, out, 0, 0, k, u.frame * 17u + static_cast<uint32_t>(iter)};
    if (queueScratch) {
        CQueues q = carve_queues(queueScratch, u.screenSize[0] * u.screenSize[1], k);
        q.merged = out;
        return by_stack(stack, [&](auto S) {
            return dispatch_c_queued<decltype(S)::value>(tile, args, q, gridCap, packSurface, s);
        });
    }
    return by_stack(stack, [&](auto S) { return dispatch_c<decltype(S)::value>(tile, args, s); });
}

// ------------------------------------------------------------------------ batched views ---
// One persistent launch covers the tiles of n views: workgroups take tiles from a global
// counter in view order (view v owns tiles [tileStart[v], tileStart[v + 1])) and run the same
// per-pixel body as the single-view kernels, so every pixel's result is bit-identical to a
// per-view launch. The table (tile prefix, then one argument record per view) sits in device
// memory; it is written on the stream by k_put_blob launches whose by-value arguments carry the
// bytes, so a captured graph holds its own copy of every table.
struct ViewTable {
    const uint32_t *tileStart;
    const void *args;
    uint32_t n;
};
constexpr int kBlobWords = 64;  // 1 KB of table per upload launch
struct Blob {
    uint4 w[kBlobWords];
};
__global__ __launch_bounds__(64) void k_put_blob(const Blob b, uint32_t nWords, uint4 *dst) {
    if (threadIdx.x < nWords) dst[threadIdx.x] = b.w[threadIdx.x];
}

enum ViewKind { kViewA = 0, kViewB = 1, kViewC = 2, kViewShade = 3 };

template <int KIND, int TX, int TY, int STACK>
__global__ __launch_bounds__(TX *TY) void k_views(const ViewTable t, uint32_t *__restrict__ counter) {
    __shared__ int32_t stack[STACK * TX * TY];
    __shared__ uint32_t sTile;
    const uint32_t total = t.tileStart[t.n];
    for (;;) {
        if (threadIdx.x == 0) sTile = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t tile = sTile;
        __syncthreads();
        if (tile >= total) break;  // uniform across the workgroup: every wave leaves together
        uint32_t v = 0;
        while (t.tileStart[v + 1] <= tile) ++v;
        const uint32_t lt = tile - t.tileStart[v];
        if constexpr (KIND == kViewShade) {
            shade_px(static_cast<const ShadeArgs *>(t.args)[v], static_cast<size_t>(lt) * (TX * TY) + threadIdx.x);
        } else {
            uint32_t tilesX;
            if constexpr (KIND == kViewA) tilesX = static_cast<const PassAArgs *>(t.args)[v].tilesX;
            else if constexpr (KIND == kViewB) tilesX = static_cast<const PassBArgs *>(t.args)[v].tilesX;
            else tilesX = static_cast<const PassCArgs *>(t.args)[v].tilesX;
            const uint32_t x = (lt % tilesX) * TX + threadIdx.x % TX;
            const uint32_t y = (lt / tilesX) * TY + threadIdx.x / TX;
            if constexpr (KIND == kViewA)
                pass_a_px<TX, TY, STACK>(static_cast<const PassAArgs *>(t.args)[v], x, y, stack + threadIdx.x);
            else if constexpr (KIND == kViewB)
                pass_b_px<TX, TY>(static_cast<const PassBArgs *>(t.args)[v], x, y);
            else
                pass_c_px<TX, TY, STACK>(static_cast<const PassCArgs *>(t.args)[v], x, y, stack + threadIdx.x);
        }
    }
}

size_t views_table_bytes(uint32_t n) { return 256 + static_cast<size_t>(n) * 512; }

// Writes the table (tile prefix + records) into `table` and launches the persistent kernel.
template <int KIND, class Args, int TX, int TY, int STACK>
static hipError_t run_views(const Args *recs, const uint32_t *tiles, uint32_t n, void *table,
                            uint32_t *counter, uint32_t gridCap, hipStream_t s) {
    static_assert(sizeof(Args) <= 512, "view record fits its table slot");
    std::vector<uint8_t> bytes(256 + sizeof(Args) * n, 0);
    uint32_t *start = reinterpret_cast<uint32_t *>(bytes.data());
    uint32_t total = 0;
    for (uint32_t v = 0; v < n; ++v) {
        start[v] = total;
        total += tiles[v];
    }
    start[n] = total;
    std::memcpy(bytes.data() + 256, recs, sizeof(Args) * n);
    const size_t words = (bytes.size() + 15) / 16;
    bytes.resize(words * 16, 0);
    for (size_t w0 = 0; w0 < words; w0 += kBlobWords) {
        Blob b{};
        const uint32_t nw = static_cast<uint32_t>(words - w0 < static_cast<size_t>(kBlobWords) ? words - w0 : kBlobWords);
        std::memcpy(b.w, bytes.data() + 16 * w0, 16 * nw);
        RK_LAUNCH("k_put_blob", k_put_blob, dim3(1), dim3(64), 0, s, b, nw, static_cast<uint4 *>(table) + w0);
    }
    RK_LAUNCH("k_zero_u32", k_zero_u32, dim3(1), dim3(64), 0, s, counter, 1u);
    const ViewTable t{static_cast<const uint32_t *>(table),
                      static_cast<const uint8_t *>(table) + 256, n};
    const uint32_t grid = total < gridCap ? total : gridCap;
    if (grid == 0) return hipGetLastError();
    RK_LAUNCH("k_views", (k_views<KIND, TX, TY, STACK>), dim3(grid), dim3(TX * TY), 0, s, t, counter);
    return hipGetLastError();
}

template <int KIND, class Args, int STACK>
static hipError_t dispatch_views(int tile, const Args *recs, const uint32_t *tiles, uint32_t n,
                                 void *table, uint32_t *counter, uint32_t gridCap, hipStream_t s) {
    switch (tile) {
        case 0: return run_views<KIND, Args, 16, 16, STACK>(recs, tiles, n, table, counter, gridCap, s);
        case 1: return run_views<KIND, Args, 64, 1, STACK>(recs, tiles, n, table, counter, gridCap, s);
        case 2: return run_views<KIND, Args, 8, 8, STACK>(recs, tiles, n, table, counter, gridCap, s);
        case 3: return run_views<KIND, Args, 16, 8, STACK>(recs, tiles, n, table, counter, gridCap, s);
        default: return run_views<KIND, Args, 32, 8, STACK>(recs, tiles, n, table, counter, gridCap, s);
    }
}

static void tiles_of(int tile, uint32_t W, uint32_t H, uint32_t &tilesX, uint32_t &nTiles) {
    const Tile t = kTiles[tile];
    tilesX = (W + t.tx - 1) / t.tx;
    nTiles = tilesX * ((H + t.ty - 1) / t.ty);
}

hipError_t launch_pass_a_views(int tile, int stack, const ViewA *views, uint32_t n, void *table,
                               uint32_t *counter, uint32_t gridCap, hipStream_t s) {
    std::vector<PassAArgs> recs(n);
    std::vector<uint32_t> tiles(n);
    for (uint32_t v = 0; v < n; ++v) {
        const ViewA &w = views[v];
        recs[v] = PassAArgs{w.u, w.sc, w.cur, w.prev, w.prevRes, w.out, 0, 0, nullptr, nullptr};
        tiles_of(tile, w.u.screenSize[0], w.u.screenSize[1], recs[v].tilesX, recs[v].nTiles);
        tiles[v] = recs[v].nTiles;
    }
    return by_stack(stack, [&](auto S) {
        return dispatch_views<kViewA, PassAArgs, decltype(S)::value>(tile, recs.data(), tiles.data(), n, table,
                                                                     counter, gridCap, s);
    });
}

hipError_t launch_pass_b_views(int tile, const ViewB *views, uint32_t n, int iter, void *table,
                               uint32_t *counter, uint32_t gridCap, hipStream_t s) {
    std::vector<PassBArgs> recs(n);
    std::vector<uint32_t> tiles(n);
    for (uint32_t v = 0; v < n; ++v) {
        const ViewB &w = views[v];
        recs[v] = PassBArgs{w.u, w.g, w.in, w.out, 0, 0, w.u.frame * 31u + static_cast<uint32_t>(iter), w.cosThr};
        tiles_of(tile, w.u.screenSize[0], w.u.screenSize[1], recs[v].tilesX, recs[v].nTiles);
        tiles[v] = recs[v].nTiles;
    }
    return dispatch_views<kViewB, PassBArgs, 1>(tile, recs.data(), tiles.data(), n, table, counter, gridCap, s);
}

hipError_t launch_pass_c_views(int tile, int stack, const ViewC *views, uint32_t n, int k, int iter,
                               void *table, uint32_t *counter, uint32_t gridCap, hipStream_t s) {
    std::vector<PassCArgs> recs(n);
    std::vector<uint32_t> tiles(n);
    for (uint32_t v = 0; v < n; ++v) {
        const ViewC &w = views[v];
        recs[v] = PassCArgs{w.u, w.sc, w.g, w.in, w.out, 0, 0, k, w.u.frame * 17u + static_cast<uint32_t>(iter)};
        tiles_of(tile, w.u.screenSize[0], w.u.screenSize[1], recs[v].tilesX, recs[v].nTiles);
        tiles[v] = recs[v].nTiles;
    }
    return by_stack(stack, [&](auto S) {
        return dispatch_views<kViewC, PassCArgs, decltype(S)::value>(tile, recs.data(), tiles.data(), n, table,
                                                                     counter, gridCap, s);
    });
}

hipError_t launch_shade_views(const ViewShade *views, uint32_t n, float gamma, void *table,
                              uint32_t *counter, uint32_t gridCap, hipStream_t s) {
    std::vector<ShadeArgs> recs(n);
    std::vector<uint32_t> tiles(n);
    for (uint32_t v = 0; v < n; ++v) {
        const ViewShade &w = views[v];
        recs[v] = ShadeArgs{w.sc, w.cam, gamma, w.g, w.res, w.rgb, w.rgbP};
        tiles[v] = static_cast<uint32_t>((static_cast<size_t>(w.g.W) * w.g.H + 255) / 256);
    }
    return run_views<kViewShade, ShadeArgs, 16, 16, 1>(recs.data(), tiles.data(), n, table, counter, gridCap, s);
}

hipError_t launch_shade(const SceneView &sc, f3 cam, float gamma, const GView &g,
                        const restir_reservoir *res, float *rgb, uint64_t rgbP, bool visibility, int stack,
                        hipStream_t s) {
    ShadeArgs args{sc, cam, gamma, g, res, rgb, rgbP};
    const size_t n = static_cast<size_t>(g.W) * g.H;
    const dim3 grid(static_cast<uint32_t>((n + 255) / 256));
    if (!visibility) {
        RK_LAUNCH("k_shade", (k_shade<24, false>), grid, dim3(256), 0, s, args);
        return hipGetLastError();
    }
    return by_stack(stack, [&](auto S) {
        RK_LAUNCH("k_shade", (k_shade<decltype(S)::value, true>), grid, dim3(256), 0, s, args);
        return hipGetLastError();
    });
}

template <int STACK, bool REF>
static void run_trace(bool persistent, uint32_t gridCap, uint32_t *counter, const TraceArgs &args,
                      hipStream_t s) {
    const uint32_t blocks = (args.n + 255) / 256;
    if (!persistent) {
        RK_LAUNCH("k_trace", (k_trace<STACK, REF>), dim3(blocks), dim3(256), 0, s, args);
        return;
    }
    RK_LAUNCH("k_zero_counters", k_zero_counters, dim3(1), dim3(64), 0, s, counter, 1u, args.sc.traceHeads);
    const uint32_t grid = blocks < gridCap ? blocks : gridCap;
    if (!REF && !args.visits && !args.tests) {
        // no counts asked for: the passes' queue tracer (gathered triangle tests)
        RK_LAUNCH("k_trace_queue(batch)", (k_trace_queue<STACK, kTraceBatch>), dim3(grid), dim3(256), 0, s, args.sc,
                  args.p1, args.p2, nullptr, args.n, CQueues{}, args.shadowed);
        return;
    }
    RK_LAUNCH("k_trace_pool", (k_trace_pool<STACK, REF>), dim3(grid), dim3(256), 0, s, args, counter);
}

hipError_t launch_trace(int stack, bool refLoop, bool persistent, uint32_t gridCap,
                        uint32_t *counter, const SceneView &sc, const float4 *p1, const float4 *p2,
                        uint32_t n, uint32_t *shadowed, uint32_t *visits, uint32_t *tests,
                        hipStream_t s) {
    TraceArgs args{sc, p1, p2, n, shadowed, visits, tests};
    return by_stack(stack, [&](auto S) {
        constexpr int kS = decltype(S)::value;
        if (refLoop) run_trace<kS, true>(persistent, gridCap, counter, args, s);
        else run_trace<kS, false>(persistent, gridCap, counter, args, s);
        return hipGetLastError();
    });
}

CamBasis make_cam_basis(const restir_camera &c, float aspect) {
    // camera.h:25-50
    CamBasis b;
    b.pos = f3{c.position[0], c.position[1], c.position[2]};
    const f3 fwd = normalize(f3{c.lookAt[0], c.lookAt[1], c.lookAt[2]} - b.pos);
    const f3 right = normalize(cross(fwd, f3{c.worldUp[0], c.worldUp[1], c.worldUp[2]}));
    const f3 up = cross(right, fwd);
    b.fwd = fwd;
    b.right = right;
    b.down = neg(up);
    const float f = 1.0f / tanf(0.5f * c.fovYRadians);
    b.sx = aspect / f;
    b.sy = 1.0f / f;
    b.A = -c.zFar / (c.zNear - c.zFar);
    b.B = c.zNear * c.zFar / (c.zNear - c.zFar);
    return b;
}

hipError_t launch_primary(int stack, const SceneView &sc, const CamBasis &cb,
                          const int32_t *triMaterial, const float *materials, float *pos,
                          void *nrm, bool nrmHalf, float *alb, float *mat, float *depth, uint64_t P, uint32_t W,
                          uint32_t H, hipStream_t s) {
    PrimaryArgs args{sc, cb, triMaterial, materials, pos, nrm, alb, mat, depth, P, W, H, nrmHalf};
    const size_t n = static_cast<size_t>(W) * H;
    const dim3 grid(static_cast<uint32_t>((n + 255) / 256));
    if (stack <= 32) RK_LAUNCH("k_primary", (k_primary<32>), grid, dim3(256), 0, s, args);
    else RK_LAUNCH("k_primary", (k_primary<64>), grid, dim3(256), 0, s, args);
    return hipGetLastError();
}

}  // namespace rk
