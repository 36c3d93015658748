// alias_build_cpu.cpp — the device alias-table build's kernel body (restir-vulkan_amd/csrc/
// alias_kernel.h, the code k_alias_build runs) on the CPU with one std::thread per lane: 1,024
// threads, a real workgroup barrier and real 64-lane wave exchanges, so that ThreadSanitizer sees
// every access between barriers and AddressSanitizer every index. With --chaos each lane yields at
// random before barriers and exchanges, to vary the interleaving. Each case's table is compared
// byte for byte with the host build (restir_build_alias_table). Test infrastructure only.
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <thread>
#include <vector>

#define AK_FN inline
#include "../../restir-vulkan_amd/csrc/alias_kernel.h"

extern "C" int restir_build_alias_table(const restir_point_light *points, uint32_t nPoint,
                                        const restir_tri_light *tris, uint32_t nTri,
                                        restir_alias_table_column *out);

namespace {

bool g_chaos = false;
const char *g_only = nullptr;  // run only the case of this name

struct CpuGroup {
    uint32_t t;
    std::barrier<> *group;  // the 1,024 lanes
    std::barrier<> *wave;   // the 64 lanes of this lane's wave
    uint32_t *slots;        // this wave's exchange slots
    std::mt19937 rng;
    uint32_t tid() const { return t; }
    void jitter() {
        if (g_chaos && (rng() & 3u) == 0) std::this_thread::yield();
    }
    void sync() {
        jitter();
        group->arrive_and_wait();
    }
    uint32_t shfl_up(uint32_t v, uint32_t d) {
        jitter();
        const uint32_t lane = t & 63u;
        slots[lane] = v;
        wave->arrive_and_wait();
        const uint32_t o = lane >= d ? slots[lane - d] : v;
        wave->arrive_and_wait();  // slots are rewritten by the next exchange
        return o;
    }
};

bool run_case(const char *name, const std::vector<restir_point_light> &pts, const std::vector<restir_tri_light> &tls,
              uint32_t seed) {
    if (g_only && std::strcmp(g_only, name) != 0) return true;
    std::printf("%s: running\n", name);
    const bool points = !pts.empty();
    const uint32_t n = static_cast<uint32_t>(points ? pts.size() : tls.size());
    std::vector<restir_alias_table_column> host(n), dev(n);
    if (restir_build_alias_table(points ? pts.data() : nullptr, points ? n : 0, points ? nullptr : tls.data(),
                                 points ? 0 : n, host.data()) != 0) {
        std::printf("%s: host build failed\n", name);
        return false;
    }
    // the device buffers: exactly n elements each, so ASan reports any index past them
    std::unique_ptr<float[]> power(new float[n]), scaled(new float[n]);
    std::unique_ptr<uint32_t[]> under(new uint32_t[n]), overA(new uint32_t[n]), overB(new uint32_t[n]);
    std::unique_ptr<ak::Shared> sh(new ak::Shared);
    std::memset(dev.data(), 0xCD, sizeof(dev[0]) * n);
    std::barrier<> group(ak::kLanes);
    std::vector<std::unique_ptr<std::barrier<>>> waves;
    for (uint32_t w = 0; w < ak::kLanes / 64; ++w) waves.emplace_back(new std::barrier<>(64));
    std::vector<uint32_t> slots(ak::kLanes);
    std::vector<std::thread> lanes;
    lanes.reserve(ak::kLanes);
    for (uint32_t t = 0; t < ak::kLanes; ++t)
        lanes.emplace_back([&, t] {
            CpuGroup g{t, &group, waves[t >> 6].get(), slots.data() + (t & ~63u), std::mt19937(seed * 7919u + t)};
            ak::build(g, *sh, points ? pts.data() : nullptr, points ? nullptr : tls.data(), n, dev.data(), power.get(),
                      scaled.get(), under.get(), overA.get(), overB.get());
        });
    for (auto &l : lanes) l.join();
    const bool same = std::memcmp(dev.data(), host.data(), sizeof(dev[0]) * n) == 0;
    std::printf("%s: %u lights, %s\n", name, n, same ? "equal to the host table" : "DIFFERS from the host table");
    return same;
}

std::vector<restir_point_light> points_of(const std::vector<float> &lum) {
    std::vector<restir_point_light> p(lum.size());
    for (size_t i = 0; i < lum.size(); ++i) {
        std::memset(&p[i], 0, sizeof p[i]);
        p[i].color_luminance[0] = p[i].color_luminance[1] = p[i].color_luminance[2] = 1.0f;
        p[i].color_luminance[3] = lum[i];
    }
    return p;
}

}  // namespace

int main(int argc, char **argv) {
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    size_t scale = 1;  // --scale N divides the case sizes (ThreadSanitizer makes each barrier slow)
    for (int i = 1; i < argc; ++i) {
        if (std::string(argv[i]) == "--chaos") g_chaos = true;
        else if (std::string(argv[i]) == "--scale" && i + 1 < argc) scale = std::strtoul(argv[++i], nullptr, 10);
        else g_only = argv[i];
    }
    if (scale < 1) scale = 1;
    std::mt19937 g(5);
    std::uniform_real_distribution<float> u01(0.0f, 1.0f);
    bool ok = true;
    const std::vector<restir_tri_light> none;
    ok &= run_case("one", points_of({2.5f}), none, 1);
    ok &= run_case("two", points_of({1.0f, 3.0f}), none, 2);
    ok &= run_case("equal", points_of(std::vector<float>(1000 / scale, 1.0f)), none, 3);
    {
        std::vector<float> l(10000 / scale);
        for (auto &v : l) v = 0.1f + 1.9f * u01(g);
        ok &= run_case("uniform_10k", points_of(l), none, 4);
    }
    {
        std::vector<float> l(5000 / scale, 1.0f);
        l.push_back(4.0e4f / static_cast<float>(scale));  // the case whose device run faulted: one large
        ok &= run_case("dominant", points_of(l), none, 5);      // paired with every small in turn
    }
    {
        std::vector<float> l(600 / scale, 1.0f);
        l.push_back(4800.0f / static_cast<float>(scale));  // the small queue runs dry while the
                                                          // dominant light is still large
        ok &= run_case("dominant_cut", points_of(l), none, 9);
    }
    {
        std::vector<float> l(3000 / scale);
        for (auto &v : l) v = u01(g) < 0.3f ? 0.0f : -std::log(1.0f - u01(g));
        ok &= run_case("with_zeros", points_of(l), none, 6);
    }
    {
        std::normal_distribution<float> nd(0.0f, 2.0f);
        std::vector<float> l(20000 / scale);
        for (auto &v : l) v = std::exp(nd(g));
        ok &= run_case("lognormal_20k", points_of(l), none, 7);
    }
    {
        std::vector<restir_tri_light> t(700 / scale + 1);
        for (auto &x : t) {
            std::memset(&x, 0, sizeof x);
            x.emission_luminance[3] = 3.0f * u01(g);
            x.normalArea[2] = 1.0f;
            x.normalArea[3] = 0.01f + u01(g);
        }
        ok &= run_case("triangle_lights", {}, t, 8);
    }
    if (g_only && std::strcmp(g_only, "lognormal_1m") == 0) {  // only when asked for by name (slow)
        std::normal_distribution<float> nd(0.0f, 2.0f);
        std::vector<float> l(1000000);
        for (auto &v : l) v = std::exp(nd(g));
        ok &= run_case("lognormal_1m", points_of(l), none, 10);
    }
    std::printf(ok ? "all cases equal to the host build\n" : "some case differs\n");
    return ok ? 0 : 1;
}
