// sanitize_check.cpp — TEST INFRASTRUCTURE ONLY. Runs the CPU oracle and the library's host
// builders (restir_build_aabb_tree, restir_build_alias_table(_mixed), restir_build_light_clusters)
// on seeded inputs, built with AddressSanitizer and UndefinedBehaviorSanitizer (make sanitize):
// an out-of-bounds access, use after free, signed overflow, misaligned or invalid shift aborts the
// program. The host builders are also compared with the oracle's restatements byte for byte.
// Cases: random triangle soups, coincident and duplicated triangles, a 2-triangle tree, alias
// tables over 1, 2, uniform, dominant (5000:1), zero-power and large lognormal weight sets, point
// plus triangle lights, k-means clusters, and a 48x36 frame of Pass A, B, C, shading and the FP64
// ground truth over a procedural room.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "restir_oracle.h"

extern "C" int restir_build_light_clusters(const restir_point_light *points, uint32_t n, uint32_t k, uint64_t seed,
                                           uint32_t iterations, int32_t *light_cluster, float *light_pwithin,
                                           int32_t *cluster_begin, int32_t *cluster_count, int32_t *cluster_lights,
                                           restir_alias_table_column *cluster_alias, float *centroids);

static int failures = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);             \
            std::fprintf(stderr, "\n");                    \
            ++failures;                                    \
        }                                                  \
    } while (0)

static std::vector<float> soup(std::mt19937 &g, uint32_t n, float extent) {
    std::uniform_real_distribution<float> u(-extent, extent);
    std::vector<float> t(9u * n);
    for (float &v : t) v = u(g);
    return t;
}

static void tree_case(const char *name, const std::vector<float> &xyz) {
    const uint32_t n = static_cast<uint32_t>(xyz.size() / 9);
    std::vector<restir_aabb_tree_node> hn(n - 1), on(n - 1);
    std::vector<restir_triangle> ht(n), ot(n);
    uint32_t hd = 0, od = 0;
    CHECK(restir_build_aabb_tree(xyz.data(), n, hn.data(), ht.data(), &hd) == 0, "%s: host build", name);
    CHECK(oracle_build_aabb_tree(xyz.data(), n, on.data(), ot.data(), &od) == 0, "%s: oracle build", name);
    CHECK(hd == od && std::memcmp(hn.data(), on.data(), hn.size() * sizeof hn[0]) == 0, "%s: trees differ", name);
    CHECK(std::memcmp(ht.data(), ot.data(), ht.size() * sizeof ht[0]) == 0, "%s: triangles differ", name);
}

static std::vector<restir_point_light> points_of(const std::vector<float> &lum, std::mt19937 &g) {
    std::uniform_real_distribution<float> u(-5.0f, 5.0f);
    std::vector<restir_point_light> p(lum.size());
    for (size_t i = 0; i < lum.size(); ++i) {
        p[i].pos[0] = u(g), p[i].pos[1] = u(g), p[i].pos[2] = u(g), p[i].pos[3] = 1.0f;
        p[i].color_luminance[0] = p[i].color_luminance[1] = p[i].color_luminance[2] = 1.0f;
        p[i].color_luminance[3] = lum[i];
    }
    return p;
}

static void alias_case(const char *name, const std::vector<float> &lum, std::mt19937 &g) {
    const auto pts = points_of(lum, g);
    const uint32_t n = static_cast<uint32_t>(lum.size());
    std::vector<restir_alias_table_column> h(n), o(n);
    CHECK(restir_build_alias_table(pts.data(), n, nullptr, 0, h.data()) == 0, "%s: host table", name);
    CHECK(oracle_build_alias_table(pts.data(), n, nullptr, 0, o.data()) == 0, "%s: oracle table", name);
    CHECK(std::memcmp(h.data(), o.data(), n * sizeof h[0]) == 0, "%s: tables differ", name);
    for (uint32_t i = 0; i < n; ++i) CHECK(h[i].alias >= 0 && h[i].alias < static_cast<int32_t>(n), "%s: alias", name);
}

int main() {
    std::mt19937 g(1234);
    // ---- trees
    tree_case("two", soup(g, 2, 1.0f));
    tree_case("soup_3000", soup(g, 3000, 10.0f));
    {
        std::vector<float> c(9 * 64);
        for (size_t i = 0; i < c.size(); ++i) c[i] = static_cast<float>(i % 9) * 0.25f;
        tree_case("coincident_64", c);
        auto d = soup(g, 40, 3.0f);
        std::vector<float> dup;
        for (int r = 0; r < 25; ++r) dup.insert(dup.end(), d.begin(), d.end());
        tree_case("duplicates_1000", dup);
    }
    // ---- alias tables
    alias_case("one", {2.5f}, g);
    alias_case("two", {1.0f, 3.0f}, g);
    {
        std::vector<float> eq(1000, 1.0f);
        alias_case("equal", eq, g);
        std::vector<float> dom(5000, 1.0f);
        dom.push_back(4.0e4f);
        alias_case("dominant", dom, g);
        std::uniform_real_distribution<float> u01(0.0f, 1.0f);
        std::exponential_distribution<float> ex(1.0f);
        std::vector<float> z(3000);
        for (float &v : z) v = u01(g) < 0.3f ? 0.0f : ex(g);
        alias_case("with_zeros", z, g);
        std::lognormal_distribution<float> ln(0.0f, 2.0f);
        std::vector<float> big(100000);
        for (float &v : big) v = ln(g);
        alias_case("lognormal_100k", big, g);
    }
    {  // point + triangle lights in one table
        std::vector<float> lum(300, 1.0f);
        const auto pts = points_of(lum, g);
        std::vector<restir_tri_light> tls(50);
        const auto tv = soup(g, 50, 2.0f);
        for (int i = 0; i < 50; ++i) {
            std::memset(&tls[i], 0, sizeof tls[i]);
            for (int c = 0; c < 3; ++c) {
                tls[i].p1[c] = tv[9 * i + c], tls[i].p2[c] = tv[9 * i + 3 + c], tls[i].p3[c] = tv[9 * i + 6 + c];
                tls[i].emission_luminance[c] = 2.0f;
            }
            tls[i].emission_luminance[3] = 2.0f;
            tls[i].normalArea[1] = 1.0f;
            tls[i].normalArea[3] = 0.5f + 0.01f * i;
        }
        std::vector<restir_alias_table_column> h(350), o(350);
        CHECK(restir_build_alias_table_mixed(pts.data(), 300, tls.data(), 50, h.data()) == 0, "mixed host");
        CHECK(oracle_build_alias_table_mixed(pts.data(), 300, tls.data(), 50, o.data()) == 0, "mixed oracle");
        CHECK(std::memcmp(h.data(), o.data(), sizeof h[0] * h.size()) == 0, "mixed tables differ");
    }
    // ---- light clusters
    {
        std::vector<float> lum(5000);
        std::uniform_real_distribution<float> u(0.01f, 10.0f);
        for (float &v : lum) v = u(g);
        const auto pts = points_of(lum, g);
        const uint32_t n = 5000, k = 256;
        std::vector<int32_t> lc(n), beg(k), cnt(k), lights(n), lc2(n), beg2(k), cnt2(k), lights2(n);
        std::vector<float> pw(n), cen(4 * k), pw2(n), cen2(4 * k);
        std::vector<restir_alias_table_column> ca(n), ca2(n);
        CHECK(restir_build_light_clusters(pts.data(), n, k, 7, 8, lc.data(), pw.data(), beg.data(), cnt.data(),
                                          lights.data(), ca.data(), cen.data()) == 0, "clusters host");
        CHECK(oracle_build_light_clusters(pts.data(), n, k, 7, 8, lc2.data(), pw2.data(), beg2.data(), cnt2.data(),
                                          lights2.data(), ca2.data(), cen2.data()) == 0, "clusters oracle");
        CHECK(lc == lc2 && lights == lights2 && std::memcmp(pw.data(), pw2.data(), 4 * n) == 0, "clusters differ");
    }
    // ---- one frame of the passes over a procedural room (floor, back wall, a box) with 64 lights
    {
        std::vector<float> xyz = {-2, 0, -2, -2, 0, 2, 2, 0, 2,   -2, 0, -2, 2, 0, 2, 2, 0, -2,
                                  -2, 0, -2, 2, 0, -2, 2, 3, -2,  -2, 0, -2, 2, 3, -2, -2, 3, -2};
        const auto box = soup(g, 40, 0.6f);
        xyz.insert(xyz.end(), box.begin(), box.end());
        const uint32_t nt = static_cast<uint32_t>(xyz.size() / 9);
        std::vector<restir_aabb_tree_node> nodes(nt - 1);
        std::vector<restir_triangle> tris(nt);
        uint32_t depth = 0;
        CHECK(restir_build_aabb_tree(xyz.data(), nt, nodes.data(), tris.data(), &depth) == 0, "room tree");
        std::vector<float> lum(64);
        std::uniform_real_distribution<float> u(0.1f, 10.0f);
        for (float &v : lum) v = u(g);
        auto pts = points_of(lum, g);
        for (auto &p : pts) p.pos[1] = std::fabs(p.pos[1]) * 0.5f + 0.5f;
        std::vector<restir_alias_table_column> alias(64);
        CHECK(restir_build_alias_table(pts.data(), 64, nullptr, 0, alias.data()) == 0, "room table");
        oracle_scene sc{nodes.data(), nt - 1, tris.data(), nt, pts.data(), 64, nullptr, 0, alias.data(), 64};
        const uint32_t W = 48, H = 36, P = W * H;
        std::vector<float> planes(13 * P);
        restir_camera cam{{0.0f, 1.5f, 4.0f}, {0.0f, 0.8f, 0.0f}, {0.0f, 1.0f, 0.0f}, 1.0471976f, 0.01f, 1000.0f};
        std::vector<int32_t> mat(nt, 0);
        std::vector<float> mats = {0.6f, 0.5f, 0.4f, 0.0f, 0.5f, 0.0f, 0.0f, 0.0f};
        float *b = planes.data();
        oracle_primary_gbuffer(tris.data(), nt, &cam, mat.data(), mats.data(), b, b + 3 * P, b + 6 * P, b + 10 * P,
                               b + 12 * P, P, W, H);
        restir_gbuffer gb{b, b + 3 * P, b + 6 * P, b + 10 * P, b + 12 * P, P, W, H, RESTIR_NORMAL_F32, 0};
        restir_uniforms uni;
        std::memset(&uni, 0, sizeof uni);
        oracle_projection_view(&cam, static_cast<float>(W) / H, uni.prevFrameProjectionViewMatrix);
        uni.cameraPos[0] = cam.position[0], uni.cameraPos[1] = cam.position[1], uni.cameraPos[2] = cam.position[2];
        uni.cameraPos[3] = 1.0f;
        uni.screenSize[0] = W, uni.screenSize[1] = H;
        uni.initialLightSampleCount = 16;
        uni.temporalSampleCountMultiplier = 20;
        uni.spatialPosThreshold = 0.1f;
        uni.spatialNormalThreshold = 25.0f;
        uni.spatialNeighbors = 5;
        uni.spatialRadius = 30.0f;
        std::vector<restir_reservoir> prev(P), a(P), bres(P), c(P);
        std::memset(prev.data(), 0, sizeof(restir_reservoir) * P);
        for (uint32_t flags : {0u, 3u, 7u}) {  // none; visibility + temporal; + standard weights
            uni.flags = flags;
            for (uint32_t f = 1; f <= 2; ++f) {
                uni.frame = f;
                oracle_pass_a(&sc, &uni, &gb, &gb, prev.data(), a.data());
                oracle_pass_b(&sc, &uni, 0, &gb, a.data(), bres.data());
                oracle_pass_c(&sc, &uni, 5, 0, &gb, a.data(), c.data());
                prev = c;
            }
        }
        std::vector<int32_t> cl(4 * P);
        std::vector<float> cq(4 * P, 0.25f);
        for (uint32_t i = 0; i < 4 * P; ++i) cl[i] = static_cast<int32_t>(i % 70) - 3;  // some out of range
        uni.initialLightSampleCount = 4;
        oracle_pass_a_candidates(&sc, &uni, &gb, &gb, prev.data(), cl.data(), cq.data(), a.data());
        std::vector<float> rgb(3 * P);
        oracle_shade(&sc, uni.cameraPos, 2.2f, &gb, c.data(), rgb.data(), P);
        std::vector<double> truth(3 * P);
        oracle_ground_truth_f64(&sc, uni.cameraPos, &gb, 1, 16, nullptr, truth.data(), P);
        double s = 0.0;
        for (double v : truth) s += v;
        CHECK(std::isfinite(s) && s > 0.0, "ground truth sum %g", s);
    }
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("sanitize_check: oracle and host builders clean under ASan/UBSan\n");
    return 0;
}
