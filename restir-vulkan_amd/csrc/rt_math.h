// rt_math.h — scalar/vector arithmetic of the ReSTIR path for gfx950 (and the host side of the
// library). Semantics are pinned so that results are reproducible bit for bit:
//   * built with -ffp-contract=off: no FMA contraction; a/b and sqrtf are IEEE correctly rounded
//     (HIP's default correctly-rounded fp32 divide/sqrt is kept);
//   * GLSL min/max/clamp/mix as the GLSL specification writes them;
//   * dot = (x*x' + y*y') + z*z'; normalize(v) = v / sqrt(dot(v, v)) per component;
//   * pow(r, 2.0) of disneyBRDF.glsl:46 is r*r;
//   * sin/cos: quadrant reduction and Taylor polynomials in double, rounded to float.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RT_HD __host__ __device__ __forceinline__

namespace rt {

constexpr float kPi = 3.14159265358979323846f;  // M_PI of disneyBRDF.glsl:1 as a float literal

struct f3 {
    float x, y, z;
};
RT_HD f3 make_f3(float x, float y, float z) { return f3{x, y, z}; }
RT_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RT_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RT_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
RT_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
RT_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
RT_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
RT_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
RT_HD f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RT_HD f3 normalize(f3 a) { return a / sqrtf(dot(a, a)); }
RT_HD f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }

RT_HD float gmin(float x, float y) { return (y < x) ? y : x; }
RT_HD float gmax(float x, float y) { return (x < y) ? y : x; }
RT_HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
RT_HD float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// ------------------------------------------------------------------------------------ PCG32
// rand.glsl:7-32 — PCG32 XSH-RR, 64-bit state, seeded as pcg32_srandom_r(seed, seq).
struct Pcg {
    uint64_t state, inc;
};
constexpr uint64_t kPcgMul = 6364136223846793005ull;
// The number drawn at state `old` (the output permutation; the state then steps to old * a + inc).
RT_HD uint32_t pcg_output(uint64_t old) {
    const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
RT_HD uint32_t pcg_next(Pcg &r) {
    const uint64_t old = r.state;
    r.state = old * kPcgMul + r.inc;
    return pcg_output(old);
}
// Jump-ahead: k steps after state s the state is a^k s + inc (1 + a + ... + a^(k-1)) (mod 2^64).
constexpr uint64_t pcg_mul_pow(int k) {
    uint64_t m = 1;
    for (int i = 0; i < k; ++i) m *= kPcgMul;
    return m;
}
constexpr uint64_t pcg_mul_sum(int k) {
    uint64_t sum = 0, m = 1;
    for (int i = 0; i < k; ++i) {
        sum += m;
        m *= kPcgMul;
    }
    return sum;
}
RT_HD Pcg pcg_seed(uint64_t seed, uint64_t seq) {
    Pcg r{0ull, (seq << 1u) | 1u};
    (void)pcg_next(r);
    r.state += seed;
    (void)pcg_next(r);
    return r;
}
// u32 -> float rounds to nearest even; the scale by 2^-32 is exact.
RT_HD float pcg_float(Pcg &r) { return static_cast<float>(pcg_next(r)) * 2.3283064365386963e-10f; }
// The float pcg_float draws at state s.
RT_HD float pcg_float_at(uint64_t s) { return static_cast<float>(pcg_output(s)) * 2.3283064365386963e-10f; }

// ---------------------------------------------------------------------------------- sin/cos
RT_HD void sincos_f(float xf, float &s_out, float &c_out) {
    const double x = static_cast<double>(xf);
    const double k = floor(x * 0.63661977236758134308 + 0.5);
    double r = x - k * 1.57079632679489661923;
    r = r - k * 6.12323399573676603587e-17;
    const double r2 = r * r;
    const double s =
        r + r * r2 *
                (-1.0 / 6.0 +
                 r2 * (1.0 / 120.0 +
                       r2 * (-1.0 / 5040.0 +
                             r2 * (1.0 / 362880.0 +
                                   r2 * (-1.0 / 39916800.0 +
                                         r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0)))))));
    const double c =
        1.0 + r2 * (-0.5 +
                    r2 * (1.0 / 24.0 +
                          r2 * (-1.0 / 720.0 +
                                r2 * (1.0 / 40320.0 +
                                      r2 * (-1.0 / 3628800.0 +
                                            r2 * (1.0 / 479001600.0 +
                                                  r2 * (-1.0 / 87178291200.0 +
                                                        r2 * (1.0 / 20922789888000.0))))))));
    const long q = static_cast<long>(k) & 3L;
    double sv, cv;
    if (q == 0) { sv = s; cv = c; }
    else if (q == 1) { sv = c; cv = -s; }
    else if (q == 2) { sv = -s; cv = -c; }
    else { sv = -c; cv = s; }
    s_out = static_cast<float>(sv);
    c_out = static_cast<float>(cv);
}
// cos(radians(deg)) of spatialReuse.comp:67
RT_HD float cos_degrees(float deg) {
    float s, c;
    sincos_f(deg * 0.0174532924f, s, c);
    return c;
}

// ------------------------------------------------------------------------------------- BRDF
RT_HD float luminance(float r, float g, float b) {  // common.glsl:7-9
    return 0.2126f * r + 0.7152f * g + 0.0722f * b;
}
RT_HD float schlick(float c) {  // disneyBRDF.glsl:5-10
    const float m = gclamp(1.0f - c, 0.0f, 1.0f);
    const float sm = m * m;
    return sm * sm * m;
}
RT_HD float gtr2(float ndoth, float a) {  // disneyBRDF.glsl:13-18
    const float a2 = a * a;
    const float t = 1.0f + (a2 - 1.0f) * ndoth * ndoth;
    return a2 / (kPi * t * t);
}
RT_HD float smith_g(float ndotv, float alphaG) {  // disneyBRDF.glsl:20-25
    const float a = alphaG * alphaG;
    const float b = ndotv * ndotv;
    return 1.0f / (fabsf(ndotv) + gmax(sqrtf(a + b - a * b), 0.0001f));
}

// Per-pixel constants of the BRDF that do not depend on the light: alpha, and the terms of
// cosOut. Hoisting them changes no rounding (same operations on the same operands).
struct Surface {
    f3 pos, n, wo;
    float cosOut, albLum, rough, metal;
    float a;          // max(0.001, rough^2)
    float fOut;       // schlick(cosOut)
    float gOut;       // smith_g(cosOut, a)
    float oneMinusMetal;
};
RT_HD Surface make_surface(f3 pos, f3 n, f3 cam, float albLum, float rough, float metal) {
    Surface s;
    s.pos = pos;
    s.n = n;
    s.wo = normalize(cam - pos);
    s.cosOut = dot(n, s.wo);
    s.albLum = albLum;
    s.rough = rough;
    s.metal = metal;
    s.a = gmax(0.001f, rough * rough);
    s.fOut = schlick(s.cosOut);
    s.gOut = smith_g(s.cosOut, s.a);
    s.oneMinusMetal = 1.0f - metal;
    return s;
}

struct LightGeom {
    float cosIn, cosHalf, cosInHalf, geometry;
    bool below;
};
// restirUtils.glsl:7-25 (shared by evaluatePHat and evaluatePHatFull)
RT_HD LightGeom light_geometry(const Surface &s, f3 L, f3 LN, bool useLN) {
    LightGeom g{};
    f3 wi = L - s.pos;
    if (dot(wi, s.n) < 0.0f) {
        g.below = true;
        return g;
    }
    const float sqrDist = dot(wi, wi);
    wi = wi / sqrtf(sqrDist);
    g.cosIn = dot(s.n, wi);
    const f3 h = normalize(wi + s.wo);
    g.cosHalf = dot(s.n, h);
    g.cosInHalf = dot(wi, h);
    g.geometry = g.cosIn / sqrDist;
    if (useLN) g.geometry *= fabsf(dot(wi, LN));
    g.below = false;
    return g;
}
// disneyBrdfDiffuseFactor (:27-33) and the specular factor Gs*Ds (:42-55)
RT_HD float diffuse_factor(const Surface &s, const LightGeom &g) {
    const float fin = schlick(g.cosIn);
    const float fd90 = 0.5f + 2.0f * g.cosInHalf * g.cosInHalf * s.rough;
    const float fd = gmix(1.0f, fd90, fin) * gmix(1.0f, fd90, s.fOut);
    return fd * s.oneMinusMetal / kPi;
}
RT_HD float gs_ds(const Surface &s, const LightGeom &g) {
    const float ds = gtr2(g.cosHalf, s.a);
    float gs = smith_g(g.cosIn, s.a);
    gs *= s.gOut;
    return gs * ds;
}
// evaluatePHat, restirUtils.glsl:3-28 with disneyBrdfLuminance (disneyBRDF.glsl:82-90)
RT_HD float eval_phat(const Surface &s, f3 L, f3 LN, bool useLN, float emLum) {
    const LightGeom g = light_geometry(s, L, LN, useLN);
    if (g.below) return 0.0f;
    float brdf = 0.0f;
    if (!(g.cosIn < 0.0f)) {
        const float diffuse = s.albLum * diffuse_factor(s, g);
        const float fih = schlick(g.cosInHalf);
        const float gsds = gs_ds(s, g);
        const float specLum = gmix(0.04f, s.albLum, s.metal);
        const float fs = gmix(specLum, 1.0f, fih);
        brdf = diffuse + fs * gsds;
    }
    return emLum * brdf * g.geometry;
}
// evaluatePHatFull, restirUtils.glsl:30-55 with disneyBrdfColor (disneyBRDF.glsl:73-81)
RT_HD f3 eval_phat_full(const Surface &s, f3 albedo, f3 L, f3 LN, bool useLN, f3 em) {
    const LightGeom g = light_geometry(s, L, LN, useLN);
    if (g.below) return f3{0.0f, 0.0f, 0.0f};
    f3 brdf{0.0f, 0.0f, 0.0f};
    if (!(g.cosIn < 0.0f)) {
        const float df = diffuse_factor(s, g);
        const float fih = schlick(g.cosInHalf);
        const float gsds = gs_ds(s, g);
        const float alb[3] = {albedo.x, albedo.y, albedo.z};
        float out[3];
        for (int c = 0; c < 3; ++c) {
            const float diffuse = alb[c] * df;
            const float sc = gmix(0.04f, alb[c], s.metal);
            const float fs = gmix(sc, 1.0f, fih);
            out[c] = diffuse + fs * gsds;
        }
        brdf = f3{out[0], out[1], out[2]};
    }
    return f3{em.x * brdf.x * g.geometry, em.y * brdf.y * g.geometry, em.z * brdf.z * g.geometry};
}

// ------------------------------------------------------------------------------- reservoirs
// Register-resident reservoir with the fields of restirStructs.glsl:19-34.
struct Res {
    f3 pos;
    float emLum;
    float4 normal;
    int32_t lightIndex;
    float pHat, sumW, w;
    uint32_t M;
};
RT_HD Res res_zero() {
    Res r;
    r.pos = f3{0.0f, 0.0f, 0.0f};
    r.emLum = 0.0f;
    r.normal = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r.lightIndex = 0;
    r.pHat = r.sumW = r.w = 0.0f;
    r.M = 0u;
    return r;
}
// updateReservoirAt, reservoir.glsl:6-26
// u: the update's uniform draw; returns true when the sample is taken
RT_HD bool res_update_u(Res &r, float weight, f3 pos, float4 n, float emLum, int32_t li, float pHat,
                        float w, float u) {
    r.sumW += weight;
    const float rp = weight / r.sumW;
    if (u < rp) {
        r.pos = pos;
        r.emLum = emLum;
        r.normal = n;
        r.lightIndex = li;
        r.pHat = pHat;
        r.w = w;
        return true;
    }
    return false;
}
RT_HD void res_update(Res &r, float weight, f3 pos, float4 n, float emLum, int32_t li, float pHat,
                      float w, Pcg &rng) {
    res_update_u(r, weight, pos, n, emLum, li, pHat, w, pcg_float(rng));
}
// addSampleToReservoir, reservoir.glsl:28-42 (u: the update's uniform draw)
RT_HD void res_add_sample_u(Res &r, f3 pos, float4 n, float emLum, int32_t li, float pHat,
                            float sampleP, float u) {
    const float weight = pHat / sampleP;
    r.M += 1u;
    const float w = (r.sumW + weight) / (static_cast<float>(r.M) * pHat);
    res_update_u(r, weight, pos, n, emLum, li, pHat, w, u);
}
RT_HD void res_add_sample(Res &r, f3 pos, float4 n, float emLum, int32_t li, float pHat,
                          float sampleP, Pcg &rng) {
    res_add_sample_u(r, pos, n, emLum, li, pHat, sampleP, pcg_float(rng));
}
// combineReservoirs, reservoir.glsl:44-64
RT_HD void res_combine(Res &self, const Res &o, float pHat, Pcg &rng) {
    self.M += o.M;
    const float weight = pHat * o.w * static_cast<float>(o.M);
    if (weight > 0.0f) res_update(self, weight, o.pos, o.normal, o.emLum, o.lightIndex, pHat, o.w, rng);
    if (self.w > 0.0f) self.w = self.sumW / (static_cast<float>(self.M) * self.pHat);
}

}  // namespace rt
