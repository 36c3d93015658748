/*
 * restir_oracle.c — CPU oracle of the ReSTIR hot path. TEST INFRASTRUCTURE ONLY: loaded by the
 * tests, the smoke check and the CPU-baseline leg of the benchmark; never shipped or measured as
 * the product. Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp, no -ffast-math).
 *
 * Every function names the reference file:line it restates (paths relative to the reference
 * repository lukedan/ReSTIR-Vulkan). Operation semantics are fixed in restir_oracle.h.
 */
#include "restir_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_F 3.14159265358979323846f

/* ------------------------------------------------------------------ GLSL scalar semantics -- */
static inline float gmin(float x, float y) { return (y < x) ? y : x; }
static inline float gmax(float x, float y) { return (x < y) ? y : x; }
static inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
static inline float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }

typedef struct { float x, y, z; } v3;
static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 smul(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 vdiv(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline float vdot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 vnorm(v3 a) { return vdiv(a, sqrtf(vdot(a, a))); }

/* ------------------------------------------------------------------------------ PCG32 ----- */
/* rand.glsl:12-18 (pcg32_random_r) */
static inline uint32_t pcg_next(uint64_t r[2]) {
    uint64_t old = r[0];
    r[0] = old * 6364136223846793005ULL + r[1];
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}
/* rand.glsl:20-28 (pcg32_srandom_r) */
static inline void pcg_seed(uint64_t r[2], uint64_t seed, uint64_t seq) {
    r[0] = 0;
    r[1] = (seq << 1u) | 1u;
    (void)pcg_next(r);
    r[0] += seed;
    (void)pcg_next(r);
}
/* rand.glsl:30-32: u32 -> float (round to nearest even), then / 2^32 (exact) */
static inline float pcg_float(uint64_t r[2]) { return (float)pcg_next(r) / 4294967296.0f; }

void oracle_pcg32_stream(uint64_t seed, uint64_t seq, uint32_t n, uint32_t *out) {
    uint64_t r[2];
    pcg_seed(r, seed, seq);
    for (uint32_t i = 0; i < n; ++i) out[i] = pcg_next(r);
}
void oracle_pcg32_float_stream(uint64_t seed, uint64_t seq, uint32_t n, float *out) {
    uint64_t r[2];
    pcg_seed(r, seed, seq);
    for (uint32_t i = 0; i < n; ++i) out[i] = pcg_float(r);
}

/* --------------------------------------------------------------------------- sin / cos ----- */
/* Quadrant reduction by k = floor(x*2/pi + 0.5) and Taylor polynomials to degree 15/16 on
 * [-pi/4, pi/4], all in double, result rounded to float. */
void oracle_sincos(float xf, float *s_out, float *c_out) {
    double x = (double)xf;
    double k = floor(x * 0.63661977236758134308 + 0.5);
    double r = x - k * 1.57079632679489661923;
    r = r - k * 6.12323399573676603587e-17;
    double r2 = r * r;
    double s = r + r * r2 * (-1.0 / 6.0 + r2 * (1.0 / 120.0 + r2 * (-1.0 / 5040.0 + r2 * (1.0 / 362880.0 +
               r2 * (-1.0 / 39916800.0 + r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0)))))));
    double c = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24.0 + r2 * (-1.0 / 720.0 + r2 * (1.0 / 40320.0 +
               r2 * (-1.0 / 3628800.0 + r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0 +
               r2 * (1.0 / 20922789888000.0))))))));
    long q = ((long)k) & 3L;
    double sv, cv;
    switch (q) {
        case 0: sv = s; cv = c; break;
        case 1: sv = c; cv = -s; break;
        case 2: sv = -s; cv = -c; break;
        default: sv = -c; cv = s; break;
    }
    *s_out = (float)sv;
    *c_out = (float)cv;
}

/* ------------------------------------------------------------------------ BRDF / p_hat ----- */
float oracle_luminance(float r, float g, float b) { /* common.glsl:7-9 */
    return 0.2126f * r + 0.7152f * g + 0.0722f * b;
}
static inline float schlick(float c) { /* disneyBRDF.glsl:5-10 */
    float m = gclamp(1.0f - c, 0.0f, 1.0f);
    float sm = m * m;
    return sm * sm * m;
}
static inline float gtr2(float ndoth, float a) { /* disneyBRDF.glsl:13-18 */
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * ndoth * ndoth;
    return a2 / (PI_F * t * t);
}
static inline float smith_g(float ndotv, float alphaG) { /* disneyBRDF.glsl:20-25 */
    float a = alphaG * alphaG;
    float b = ndotv * ndotv;
    return 1.0f / (fabsf(ndotv) + gmax(sqrtf(a + b - a * b), 0.0001f));
}
static inline float diffuse_factor(float ci, float co, float cih, float rough, float metal) {
    /* disneyBRDF.glsl:27-33 */
    float fin = schlick(ci);
    float fout = schlick(co);
    float fd90 = 0.5f + 2.0f * cih * cih * rough;
    float fd = gmix(1.0f, fd90, fin) * gmix(1.0f, fd90, fout);
    return fd * (1.0f - metal) / PI_F;
}
static inline void spec_factors(float ci, float co, float ch, float cih, float rough, float *fih,
                                float *gsds) { /* disneyBRDF.glsl:42-55, pow(r,2.0) = r*r */
    *fih = schlick(cih);
    float a = gmax(0.001f, rough * rough);
    float ds = gtr2(ch, a);
    float gs = smith_g(ci, a);
    gs *= smith_g(co, a);
    *gsds = gs * ds;
}
float oracle_brdf_luminance(float ci, float co, float ch, float cih, float lum, float rough,
                            float metal) { /* disneyBRDF.glsl:82-90 */
    if (ci < 0.0f) return 0.0f;
    float diffuse = lum * diffuse_factor(ci, co, cih, rough, metal);
    float fih, gsds;
    spec_factors(ci, co, ch, cih, rough, &fih, &gsds);
    float specLum = gmix(0.04f, lum, metal);
    float fs = gmix(specLum, 1.0f, fih);
    float specular = fs * gsds;
    return diffuse + specular;
}
void oracle_brdf_color(float ci, float co, float ch, float cih, const float albedo[3], float rough,
                       float metal, float out[3]) { /* disneyBRDF.glsl:73-81 */
    if (ci < 0.0f) { out[0] = out[1] = out[2] = 0.0f; return; }
    float df = diffuse_factor(ci, co, cih, rough, metal);
    float fih, gsds;
    spec_factors(ci, co, ch, cih, rough, &fih, &gsds);
    for (int c = 0; c < 3; ++c) {
        float diffuse = albedo[c] * df;
        float specColor = gmix(0.04f, albedo[c], metal);
        float fs = gmix(specColor, 1.0f, fih);
        out[c] = diffuse + fs * gsds;
    }
}

typedef struct {
    float cosIn, cosOut, cosHalf, cosInHalf, geometry;
    int below;
} shading_geom;

static inline shading_geom phat_geometry(v3 P, v3 L, v3 C, v3 N, v3 LN, int useLN) {
    /* restirUtils.glsl:7-25 */
    shading_geom g;
    memset(&g, 0, sizeof g);
    v3 wi = vsub(L, P);
    if (vdot(wi, N) < 0.0f) { g.below = 1; return g; }
    float sqrDist = vdot(wi, wi);
    wi = vdiv(wi, sqrtf(sqrDist));
    v3 wo = vnorm(vsub(C, P));
    g.cosIn = vdot(N, wi);
    g.cosOut = vdot(N, wo);
    v3 h = vnorm(vadd(wi, wo));
    g.cosHalf = vdot(N, h);
    g.cosInHalf = vdot(wi, h);
    g.geometry = g.cosIn / sqrDist;
    if (useLN) g.geometry *= fabsf(vdot(wi, LN));
    return g;
}
static inline float phat_v(v3 P, v3 L, v3 C, v3 N, v3 LN, int useLN, float albLum, float emLum,
                           float rough, float metal) { /* restirUtils.glsl:3-28 */
    shading_geom g = phat_geometry(P, L, C, N, LN, useLN);
    if (g.below) return 0.0f;
    return emLum * oracle_brdf_luminance(g.cosIn, g.cosOut, g.cosHalf, g.cosInHalf, albLum, rough,
                                         metal) * g.geometry;
}
float oracle_phat(const float P[3], const float L[3], const float C[3], const float N[3],
                  const float LN[3], int useLN, float albLum, float emLum, float rough,
                  float metal) {
    return phat_v(ld3(P), ld3(L), ld3(C), ld3(N), ld3(LN), useLN, albLum, emLum, rough, metal);
}
static inline void phat_full_v(v3 P, v3 L, v3 C, v3 N, v3 LN, int useLN, const float albedo[3],
                               const float em[3], float rough, float metal, float out[3]) {
    /* restirUtils.glsl:30-55 */
    shading_geom g = phat_geometry(P, L, C, N, LN, useLN);
    if (g.below) { out[0] = out[1] = out[2] = 0.0f; return; }
    float brdf[3];
    oracle_brdf_color(g.cosIn, g.cosOut, g.cosHalf, g.cosInHalf, albedo, rough, metal, brdf);
    for (int c = 0; c < 3; ++c) out[c] = em[c] * brdf[c] * g.geometry;
}
void oracle_phat_full(const float P[3], const float L[3], const float C[3], const float N[3],
                      const float LN[3], int useLN, const float albedo[3], const float em[3],
                      float rough, float metal, float out[3]) {
    phat_full_v(ld3(P), ld3(L), ld3(C), ld3(N), ld3(LN), useLN, albedo, em, rough, metal, out);
}

/* FP64 restatement of the same integrand (luminance form), used by tests as an independent
 * closed-form check of the FP32 arithmetic. */
static double schlick_d(double c) {
    double m = c < 0.0 ? 1.0 : (c > 1.0 ? 0.0 : 1.0 - c);
    return m * m * m * m * m;
}
static void brdf_parts_d(double ci, double co, double ch, double cih, double rough, double metal,
                         double *diff_factor, double *fih, double *gsds) {
    double fd90 = 0.5 + 2.0 * cih * cih * rough;
    double fin = schlick_d(ci), fout = schlick_d(co);
    double fd = ((1.0 - fin) + fd90 * fin) * ((1.0 - fout) + fd90 * fout);
    *diff_factor = fd * (1.0 - metal) / M_PI;
    *fih = schlick_d(cih);
    double a = rough * rough;
    if (a < 0.001) a = 0.001;
    double a2 = a * a;
    double t = 1.0 + (a2 - 1.0) * ch * ch;
    double ds = a2 / (M_PI * t * t);
    double ga = a * a;
    double gi = 1.0 / (fabs(ci) + fmax(sqrt(ga + ci * ci - ga * ci * ci), 0.0001));
    double go = 1.0 / (fabs(co) + fmax(sqrt(ga + co * co - ga * co * co), 0.0001));
    *gsds = gi * go * ds;
}
double oracle_phat_f64(const double P[3], const double L[3], const double C[3], const double N[3],
                       double albLum, double emLum, double rough, double metal) {
    double wi[3] = {L[0] - P[0], L[1] - P[1], L[2] - P[2]};
    double dn = wi[0] * N[0] + wi[1] * N[1] + wi[2] * N[2];
    if (dn < 0.0) return 0.0;
    double d2 = wi[0] * wi[0] + wi[1] * wi[1] + wi[2] * wi[2];
    double d = sqrt(d2);
    for (int i = 0; i < 3; ++i) wi[i] /= d;
    double wo[3] = {C[0] - P[0], C[1] - P[1], C[2] - P[2]};
    double lo = sqrt(wo[0] * wo[0] + wo[1] * wo[1] + wo[2] * wo[2]);
    for (int i = 0; i < 3; ++i) wo[i] /= lo;
    double h[3] = {wi[0] + wo[0], wi[1] + wo[1], wi[2] + wo[2]};
    double lh = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    for (int i = 0; i < 3; ++i) h[i] /= lh;
    double ci = N[0] * wi[0] + N[1] * wi[1] + N[2] * wi[2];
    double co = N[0] * wo[0] + N[1] * wo[1] + N[2] * wo[2];
    double ch = N[0] * h[0] + N[1] * h[1] + N[2] * h[2];
    double cih = wi[0] * h[0] + wi[1] * h[1] + wi[2] * h[2];
    if (ci < 0.0) return 0.0;
    double df, fih, gsds;
    brdf_parts_d(ci, co, ch, cih, rough, metal, &df, &fih, &gsds);
    double specLum = 0.04 * (1.0 - metal) + albLum * metal;
    double fs = specLum * (1.0 - fih) + fih;
    return emLum * (albLum * df + fs * gsds) * (ci / d2);
}

/* ------------------------------------------------------------------------- reservoirs ----- */
static inline void update_at(restir_reservoir *r, float weight, v3 pos, const float n4[4],
                             float emLum, int32_t li, float pHat, float w, uint64_t rng[2]) {
    /* reservoir.glsl:6-26 */
    restir_light_sample *s = &r->samples[0];
    s->sumWeights += weight;
    float rp = weight / s->sumWeights;
    if (pcg_float(rng) < rp) {
        s->position_emissionLum[0] = pos.x;
        s->position_emissionLum[1] = pos.y;
        s->position_emissionLum[2] = pos.z;
        s->position_emissionLum[3] = emLum;
        s->normal[0] = n4[0];
        s->normal[1] = n4[1];
        s->normal[2] = n4[2];
        s->normal[3] = n4[3];
        s->lightIndex = li;
        s->pHat = pHat;
        s->w = w;
    }
}
void oracle_add_sample(restir_reservoir *res, const float position[3], const float normal4[4],
                       float emLum, int32_t li, float pHat, float sampleP, uint64_t rng[2]) {
    /* reservoir.glsl:28-42 */
    float weight = pHat / sampleP;
    res->numStreamSamples += 1u;
    float w = (res->samples[0].sumWeights + weight) / ((float)res->numStreamSamples * pHat);
    update_at(res, weight, ld3(position), normal4, emLum, li, pHat, w, rng);
}
void oracle_combine(restir_reservoir *self, const restir_reservoir *other, float pHat,
                    uint64_t rng[2]) {
    /* reservoir.glsl:44-64 */
    self->numStreamSamples += other->numStreamSamples;
    const restir_light_sample *o = &other->samples[0];
    float weight = pHat * o->w * (float)other->numStreamSamples;
    if (weight > 0.0f) {
        update_at(self, weight, ld3(o->position_emissionLum), o->normal,
                  o->position_emissionLum[3], o->lightIndex, pHat, o->w, rng);
    }
    restir_light_sample *s = &self->samples[0];
    if (s->w > 0.0f) s->w = s->sumWeights / ((float)self->numStreamSamples * s->pHat);
}

/* ------------------------------------------------------------------------ alias table ----- */
/* misc.cpp:418-497 — Vose's method with two FIFO queues over weights w[0..n) (modified) */
static void vose_fifo(float *p, uint32_t n, restir_alias_table_column *out) {
    uint32_t *big = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint32_t *small = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    size_t bh = 0, bt = 0, sh = 0, st = 0, cap = n + 1;
    float powerSum = 0.0f;
    for (uint32_t i = 0; i < n; ++i) powerSum += p[i];
    for (uint32_t i = 0; i < n; ++i) {
        out[i].prob = 0.0f;
        out[i].alias = -1;
        out[i].oriProb = p[i] / powerSum;
        out[i].aliasOriProb = 0.0f;
        p[i] = (float)n * p[i] / powerSum;
        if (p[i] >= 1.0f) big[bt++ % cap] = i; else small[st++ % cap] = i;
    }
    while (bh != bt && sh != st) {
        uint32_t g = big[bh++ % cap];
        uint32_t l = small[sh++ % cap];
        out[l].prob = p[l];
        out[l].alias = (int32_t)g;
        p[g] = (p[g] + p[l]) - 1.0f;
        if (p[g] < 1.0f) small[st++ % cap] = g; else big[bt++ % cap] = g;
    }
    while (bh != bt) { uint32_t g = big[bh++ % cap]; out[g].prob = 1.0f; out[g].alias = (int32_t)g; }
    while (sh != st) { uint32_t l = small[sh++ % cap]; out[l].prob = 1.0f; out[l].alias = (int32_t)l; }
    for (uint32_t i = 0; i < n; ++i) out[i].aliasOriProb = out[out[i].alias].oriProb;
    free(big); free(small);
}

int oracle_build_alias_table(const restir_point_light *points, uint32_t n_point,
                             const restir_tri_light *tris, uint32_t n_tri,
                             restir_alias_table_column *out) {
    /* misc.cpp:418-497: point lights if any, else triangle lights */
    uint32_t n = n_point > 0 ? n_point : n_tri;
    if (n == 0) return 0;
    float *p = (float *)malloc(sizeof(float) * n);
    for (uint32_t i = 0; i < n; ++i)
        p[i] = n_point > 0 ? points[i].color_luminance[3]
                           : tris[i].emission_luminance[3] * tris[i].normalArea[3];
    vose_fifo(p, n, out);
    free(p);
    return 0;
}

int oracle_build_alias_table_mixed(const restir_point_light *points, uint32_t n_point,
                                   const restir_tri_light *tris, uint32_t n_tri,
                                   restir_alias_table_column *out) {
    /* extension: one table over point lights then triangle lights, weights as misc.cpp:426-442 */
    uint32_t n = n_point + n_tri;
    if (n == 0) return 0;
    float *p = (float *)malloc(sizeof(float) * n);
    for (uint32_t i = 0; i < n_point; ++i) p[i] = points[i].color_luminance[3];
    for (uint32_t i = 0; i < n_tri; ++i)
        p[n_point + i] = tris[i].emission_luminance[3] * tris[i].normalArea[3];
    vose_fifo(p, n, out);
    free(p);
    return 0;
}

/* ------------------------------------------------------------------------- BVH build ------ */
/* nvmath.inl:2414-2425 semantics */
static inline float nvmin(float a, float b) { return (a < b) ? a : b; }
static inline float nvmax(float a, float b) { return (a > b) ? a : b; }

typedef struct { float c[3], mn[3], mx[3]; int32_t geom; size_t bucket; } leaf_t;
typedef struct { float mn[3], mx[3]; size_t count; } bucket_t;
typedef struct { int64_t parent; int side; size_t beg, end; } step_t; /* parent -1: dummy root */

static inline float sah(const float mn[3], const float mx[3]) { /* aabbTreeBuilder.cpp:15-18 */
    float sx = mx[0] - mn[0], sy = mx[1] - mn[1], sz = mx[2] - mn[2];
    return sx * sy + sx * sz + sy * sz;
}
static inline void bucket_init(bucket_t *b) {
    for (int k = 0; k < 3; ++k) { b->mn[k] = FLT_MAX; b->mx[k] = -FLT_MAX; }
    b->count = 0;
}
static inline float bucket_h(const bucket_t *b) { return (float)b->count * sah(b->mn, b->mx); }
static inline bucket_t bucket_merge(bucket_t l, const bucket_t *r) {
    l.count += r->count;
    for (int k = 0; k < 3; ++k) { l.mn[k] = nvmin(l.mn[k], r->mn[k]); l.mx[k] = nvmax(l.mx[k], r->mx[k]); }
    return l;
}
static inline void set_box(float dst[4], const float src[3]) {
    dst[0] = src[0]; dst[1] = src[1]; dst[2] = src[2]; dst[3] = 1.0f;
}

int oracle_build_aabb_tree(const float *xyz, uint32_t n, restir_aabb_tree_node *nodes,
                           restir_triangle *tris, uint32_t *depth_out) {
    /* aabbTreeBuilder.cpp:52-214 */
    if (n < 2) return -1;
    leaf_t *leaves = (leaf_t *)malloc(sizeof(leaf_t) * n);
    for (uint32_t i = 0; i < n; ++i) {
        const float *t = xyz + 9 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            tris[i].p1[k] = t[k]; tris[i].p2[k] = t[3 + k]; tris[i].p3[k] = t[6 + k];
        }
        tris[i].p1[3] = tris[i].p2[3] = tris[i].p3[3] = 1.0f;
        leaf_t *L = &leaves[i];
        for (int k = 0; k < 3; ++k) { /* aabbForTriangle, :8-14 */
            float a = t[k], b = t[3 + k], c = t[6 + k];
            float mn = a, mx = a;
            mn = nvmin(mn, b); mx = nvmax(mx, b);
            mn = nvmin(mn, c); mx = nvmax(mx, c);
            L->mn[k] = mn; L->mx[k] = mx;
            L->c[k] = 0.5f * (mn + mx);
        }
        L->geom = (int32_t)i;
        L->bucket = 0;
    }
    memset(nodes, 0, sizeof(restir_aabb_tree_node) * (n - 1));
    uint32_t *depth = (uint32_t *)calloc(n, sizeof(uint32_t));
    step_t *q = (step_t *)malloc(sizeof(step_t) * (2 * (size_t)n + 2));
    size_t qh = 0, qt = 0;
    int32_t alloc = 0, dummy = -1;
    uint32_t maxDepth = 0;
    q[qt].parent = -1; q[qt].side = 0; q[qt].beg = 0; q[qt].end = n; ++qt;
    while (qh != qt) {
        step_t st = q[qh++];
        int32_t *parentPtr = st.parent < 0 ? &dummy
                             : (st.side == 0 ? &nodes[st.parent].leftChild : &nodes[st.parent].rightChild);
        uint32_t myDepth = st.parent < 0 ? 1 : depth[st.parent] + 1;
        size_t len = st.end - st.beg;
        if (len == 1) {
            *parentPtr = ~leaves[st.beg].geom;
        } else if (len == 2) {
            leaf_t *l = &leaves[st.beg], *r = &leaves[st.beg + 1];
            int32_t ni = alloc++;
            *parentPtr = ni;
            depth[ni] = myDepth;
            if (myDepth > maxDepth) maxDepth = myDepth;
            nodes[ni].leftChild = ~l->geom;
            nodes[ni].rightChild = ~r->geom;
            set_box(nodes[ni].leftAabbMin, l->mn); set_box(nodes[ni].leftAabbMax, l->mx);
            set_box(nodes[ni].rightAabbMin, r->mn); set_box(nodes[ni].rightAabbMax, r->mx);
        } else {
            float cmin[3], cmax[3], amin[3], amax[3];
            for (int k = 0; k < 3; ++k) {
                cmin[k] = cmax[k] = leaves[st.beg].c[k];
                amin[k] = leaves[st.beg].mn[k]; amax[k] = leaves[st.beg].mx[k];
            }
            for (size_t i = st.beg + 1; i < st.end; ++i) {
                for (int k = 0; k < 3; ++k) {
                    cmin[k] = nvmin(cmin[k], leaves[i].c[k]);
                    cmax[k] = nvmax(cmax[k], leaves[i].c[k]);
                    amin[k] = nvmin(amin[k], leaves[i].mn[k]);
                    amax[k] = nvmax(amax[k], leaves[i].mx[k]);
                }
            }
            float outer = sah(amin, amax);
            float span[3] = {cmax[0] - cmin[0], cmax[1] - cmin[1], cmax[2] - cmin[2]};
            int d = span[0] > span[1] ? 0 : 1;
            if (span[2] > span[d]) d = 2;
            bucket_t b[12];
            for (int j = 0; j < 12; ++j) bucket_init(&b[j]);
            float range = span[d] / 12.0f;
            for (size_t i = st.beg; i < st.end; ++i) {
                float v = (leaves[i].c[d] - cmin[d]) / range;
                v = (v < 0.5f) ? 0.5f : v; /* nv_clamp */
                v = (v > 11.5f) ? 11.5f : v;
                /* NaN (all centroids coincide) is undefined behaviour in the reference; bucket 0 */
                size_t bi = (v == v) ? (size_t)v : 0;
                leaves[i].bucket = bi;
                for (int k = 0; k < 3; ++k) {
                    b[bi].mn[k] = nvmin(b[bi].mn[k], leaves[i].mn[k]);
                    b[bi].mx[k] = nvmax(b[bi].mx[k], leaves[i].mx[k]);
                }
                b[bi].count++;
            }
            bucket_t cache[11];
            {
                bucket_t cur = b[11];
                for (size_t i = 11; i > 0;) {
                    cache[--i] = cur;
                    cur = bucket_merge(cur, &b[i]);
                }
            }
            size_t opt = 0;
            float lmin[3] = {0, 0, 0}, lmax[3] = {0, 0, 0}, rmin[3] = {0, 0, 0}, rmax[3] = {0, 0, 0};
            {
                float minH = FLT_MAX;
                bucket_t sumL;
                bucket_init(&sumL);
                for (size_t sp = 0; sp < 11; ++sp) {
                    sumL = bucket_merge(sumL, &b[sp]);
                    const bucket_t *sumR = &cache[sp];
                    float h = 0.125f + (bucket_h(&sumL) + bucket_h(sumR)) / outer;
                    if (h < minH) {
                        minH = h;
                        opt = sp;
                        memcpy(lmin, sumL.mn, sizeof lmin); memcpy(lmax, sumL.mx, sizeof lmax);
                        memcpy(rmin, sumR->mn, sizeof rmin); memcpy(rmax, sumR->mx, sizeof rmax);
                    }
                }
            }
            size_t pivot = st.beg;
            for (size_t i = st.beg; i < st.end; ++i) {
                if (leaves[i].bucket <= opt) {
                    leaf_t tmp = leaves[i]; leaves[i] = leaves[pivot]; leaves[pivot] = tmp;
                    ++pivot;
                }
            }
            if (pivot == st.beg || pivot == st.end) {
                pivot = (st.beg + st.end) / 2;
                memcpy(lmin, leaves[st.beg].mn, sizeof lmin); memcpy(lmax, leaves[st.beg].mx, sizeof lmax);
                for (size_t i = st.beg + 1; i < pivot; ++i)
                    for (int k = 0; k < 3; ++k) {
                        lmin[k] = nvmin(lmin[k], leaves[i].mn[k]);
                        lmax[k] = nvmax(lmax[k], leaves[i].mx[k]);
                    }
                memcpy(rmin, leaves[pivot].mn, sizeof rmin); memcpy(rmax, leaves[pivot].mx, sizeof rmax);
                for (size_t i = pivot; i < st.end; ++i)
                    for (int k = 0; k < 3; ++k) {
                        rmin[k] = nvmin(rmin[k], leaves[i].mn[k]);
                        rmax[k] = nvmax(rmax[k], leaves[i].mx[k]);
                    }
            }
            int32_t ni = alloc++;
            *parentPtr = ni;
            depth[ni] = myDepth;
            if (myDepth > maxDepth) maxDepth = myDepth;
            set_box(nodes[ni].leftAabbMin, lmin); set_box(nodes[ni].leftAabbMax, lmax);
            set_box(nodes[ni].rightAabbMin, rmin); set_box(nodes[ni].rightAabbMax, rmax);
            q[qt].parent = ni; q[qt].side = 0; q[qt].beg = st.beg; q[qt].end = pivot; ++qt;
            q[qt].parent = ni; q[qt].side = 1; q[qt].beg = pivot; q[qt].end = st.end; ++qt;
        }
    }
    free(q); free(leaves); free(depth);
    if (depth_out) *depth_out = maxDepth;
    return (dummy == 0 && alloc == (int32_t)(n - 1)) ? 0 : -2;
}

/* ------------------------------------------------------------------------- traversal ------ */
static inline int ray_aabb_v(v3 o, v3 d, const float *mn4, const float *mx4) {
    /* softwareRaytracing.glsl:9-14 */
    v3 a = vsub(ld3(mn4), o), b = vsub(ld3(mx4), o);
    a = mk(a.x / d.x, a.y / d.y, a.z / d.z);
    b = mk(b.x / d.x, b.y / d.y, b.z / d.z);
    v3 lo = mk(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z));
    v3 hi = mk(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z));
    float rmin = gmax(lo.x, gmax(lo.y, lo.z));
    float rmax = gmin(hi.x, gmin(hi.y, hi.z));
    return rmin < 1.0f && rmax >= rmin && rmax > 0.0f;
}
int oracle_ray_aabb(const float o[3], const float d[3], const float mn[3], const float mx[3]) {
    return ray_aabb_v(ld3(o), ld3(d), mn, mx);
}
static inline int ray_tri_v(const restir_triangle *t, v3 o, v3 d) {
    /* softwareRaytracing.glsl:15-37 */
    v3 p1 = ld3(t->p1);
    v3 e1 = vsub(ld3(t->p2), p1);
    v3 e2 = vsub(ld3(t->p3), p1);
    v3 p = vcross(d, e2);
    float f = 1.0f / vdot(e1, p);
    v3 s = vsub(o, p1);
    float bx = f * vdot(s, p);
    if (bx < 0.0f || bx > 1.0f) return 0;
    v3 q = vcross(s, e1);
    float by = f * vdot(d, q);
    if (by < 0.0f || by + bx > 1.0f) return 0;
    f *= vdot(e2, q);
    return f > 0.0f && f < 1.0f;
}
int oracle_ray_triangle(const restir_triangle *t, const float o[3], const float d[3]) {
    return ray_tri_v(t, ld3(o), ld3(d));
}

#define ORACLE_STACK 256
/* softwareRaytracing.glsl:42-85. Returns 1 if unoccluded. The reference's stack holds 32 entries
 * (:40); the oracle allows deeper trees and flags overflow through *overflow. */
static int raytrace_v(const oracle_scene *s, v3 o, v3 d, uint32_t *visits, uint32_t *tests,
                      int *overflow) {
    int32_t stack[ORACLE_STACK];
    int top = 1;
    stack[0] = 0;
    int32_t cand[16];
    int nc = 0, counter = 0;
    uint32_t nv = 0, nt = 0;
    int result = 1;
    while (top > 0) {
        const restir_aabb_tree_node *node = &s->nodes[stack[--top]];
        ++nv;
        int li = ray_aabb_v(o, d, node->leftAabbMin, node->leftAabbMax);
        int ri = ray_aabb_v(o, d, node->rightAabbMin, node->rightAabbMax);
        if (li) {
            if (node->leftChild < 0) cand[nc++] = ~node->leftChild;
            else stack[top++] = node->leftChild;
        }
        if (ri) {
            if (node->rightChild < 0) cand[nc++] = ~node->rightChild;
            else stack[top++] = node->rightChild;
        }
        if (top > 32 && overflow) *overflow = 1;
        if (++counter == 8) {
            for (int i = 0; i < nc; ++i) {
                ++nt;
                if (ray_tri_v(&s->tris[cand[i]], o, d)) { result = 0; goto done; }
            }
            nc = 0;
            counter = 0;
        }
    }
    for (int i = 0; i < nc; ++i) {
        ++nt;
        if (ray_tri_v(&s->tris[cand[i]], o, d)) { result = 0; goto done; }
    }
done:
    if (visits) *visits = nv;
    if (tests) *tests = nt;
    return result;
}
/* visibilityTest.glsl:1-30: shrink the segment by tMin = 0.001 at both ends */
static inline void shadow_segment(v3 p1, v3 p2, v3 *o, v3 *d) {
    float tMin = 0.001f;
    v3 dir = vsub(p2, p1);
    v3 offset = smul(tMin, vnorm(dir));
    *o = vadd(p1, offset);
    *d = vsub(dir, smul(2.0f, offset));
}
static inline int test_visibility(const oracle_scene *s, v3 p1, v3 p2) { /* returns shadowed */
    v3 o, d;
    shadow_segment(p1, p2, &o, &d);
    return !raytrace_v(s, o, d, NULL, NULL, NULL);
}
void oracle_trace_shadow(const oracle_scene *s, const float *p1, const float *p2, uint32_t n,
                         uint32_t *shadowed, uint32_t *visits, uint32_t *tests) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        v3 o, d;
        shadow_segment(ld3(p1 + 4 * i), ld3(p2 + 4 * i), &o, &d);
        uint32_t nv = 0, nt = 0;
        int vis = raytrace_v(s, o, d, &nv, &nt, NULL);
        shadowed[i] = vis ? 0u : 1u;
        if (visits) visits[i] = nv;
        if (tests) tests[i] = nt;
    }
}

/* FP64 segment/triangle test; returns 1 on a hit with t in (0,1) and barycentrics in [0,1]. */
static int seg_tri_f64(const restir_triangle *t, const double o[3], const double d[3],
                       double *edge) {
    double e1[3], e2[3], s[3], p[3], q[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = (double)t->p2[k] - t->p1[k];
        e2[k] = (double)t->p3[k] - t->p1[k];
        s[k] = o[k] - t->p1[k];
    }
    p[0] = d[1] * e2[2] - d[2] * e2[1]; p[1] = d[2] * e2[0] - d[0] * e2[2]; p[2] = d[0] * e2[1] - d[1] * e2[0];
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0) return 0;
    double f = 1.0 / det;
    double u = f * (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]);
    q[0] = s[1] * e1[2] - s[2] * e1[1]; q[1] = s[2] * e1[0] - s[0] * e1[2]; q[2] = s[0] * e1[1] - s[1] * e1[0];
    double v = f * (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]);
    double tt = f * (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]);
    double w = 1.0 - u - v;
    double m = fmin(fabs(u), fmin(fabs(v), fabs(w)));
    double mt = fmin(fabs(tt), fabs(1.0 - tt));
    if (edge) {
        /* distance of the crossing to the triangle boundary or to a segment end, for crossings
         * that are near-inside; used to exclude grazing rays from bit-exact comparisons */
        if (u > -1e-3 && v > -1e-3 && w > -1e-3 && tt > -1e-3 && tt < 1.0 + 1e-3) {
            double e = fmin(m, mt);
            if (e < *edge) *edge = e;
        }
    }
    return u >= 0.0 && v >= 0.0 && w >= 0.0 && tt > 0.0 && tt < 1.0;
}
void oracle_trace_brute_f64(const restir_triangle *tris, uint32_t n_tris, const float *p1,
                            const float *p2, uint32_t n, uint32_t *shadowed, double *edge_dist) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        double a[3], b[3], d[3], o[3];
        for (int k = 0; k < 3; ++k) { a[k] = p1[4 * i + k]; b[k] = p2[4 * i + k]; d[k] = b[k] - a[k]; }
        double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int k = 0; k < 3; ++k) {
            double off = 0.001 * d[k] / len;
            o[k] = a[k] + off;
            d[k] = d[k] - 2.0 * off;
        }
        double e = 1.0;
        uint32_t hit = 0;
        for (uint32_t t = 0; t < n_tris; ++t) {
            if (seg_tri_f64(&tris[t], o, d, edge_dist ? &e : NULL)) {
                hit = 1;
                if (!edge_dist) break;
            }
        }
        shadowed[i] = hit;
        if (edge_dist) edge_dist[i] = e;
    }
}
static uint32_t count_rec(const oracle_scene *s, int32_t node, v3 o, v3 d) {
    const restir_aabb_tree_node *nd = &s->nodes[node];
    uint32_t c = 1;
    if (ray_aabb_v(o, d, nd->leftAabbMin, nd->leftAabbMax) && nd->leftChild >= 0)
        c += count_rec(s, nd->leftChild, o, d);
    if (ray_aabb_v(o, d, nd->rightAabbMin, nd->rightAabbMax) && nd->rightChild >= 0)
        c += count_rec(s, nd->rightChild, o, d);
    return c;
}
void oracle_count_reachable(const oracle_scene *s, const float *p1, const float *p2, uint32_t n,
                            uint32_t *count) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        v3 o, d;
        shadow_segment(ld3(p1 + 4 * i), ld3(p2 + 4 * i), &o, &d);
        count[i] = count_rec(s, 0, o, d);
    }
}

/* --------------------------------------------------------------------------- passes ------- */
typedef struct { v3 pos, n, alb; float rough, metal, depth, emissive; } texel;
static inline texel fetch(const restir_gbuffer *g, uint32_t x, uint32_t y) {
    size_t i = (size_t)y * g->width + x, P = g->plane_stride;
    texel t;
    t.pos = mk(g->position[i], g->position[P + i], g->position[2 * P + i]);
    t.n = mk(((const float *)g->normal)[i], ((const float *)g->normal)[P + i], ((const float *)g->normal)[2 * P + i]);
    t.alb = mk(g->albedo[i], g->albedo[P + i], g->albedo[2 * P + i]);
    t.emissive = g->albedo[3 * P + i];
    t.rough = g->material[i];
    t.metal = g->material[P + i];
    t.depth = g->depth ? g->depth[i] : 0.0f;
    return t;
}
static inline v3 res_pos(const restir_reservoir *r) { return ld3(r->samples[0].position_emissionLum); }
static inline float res_phat_at(const restir_reservoir *r, v3 P, v3 C, v3 N, float albLum,
                                float rough, float metal) {
    const restir_light_sample *s = &r->samples[0];
    return phat_v(P, res_pos(r), C, N, ld3(s->normal), s->normal[3] > 0.5f, albLum,
                  s->position_emissionLum[3], rough, metal);
}

/* Pass A (restirOmni.glsl:86-212). With cand_light == NULL the M candidates come from the alias
 * table as in the reference; otherwise candidate i of pixel p is (cand_light, cand_q)[i*W*H + p]:
 * a light of the alias table's index space (point lights first when the table covers both kinds)
 * drawn with probability cand_q, weighted p_hat / q in the same reservoir update
 * (addSampleToReservoir, reservoir.glsl:28-42). A candidate outside [0, n_lights) or with q <= 0
 * is empty: p_hat 0, probability 1, still counted in M and still consuming the update's draw. A
 * triangle-light candidate draws its point as the reference does (two draws, restirOmni.glsl:
 * 122-131) and divides q by |cos| * area. */
static void pass_a_impl(const oracle_scene *s, const restir_uniforms *u, const restir_gbuffer *cur,
                        const restir_gbuffer *prev, const restir_reservoir *prev_res,
                        const int32_t *cand_light, const float *cand_q, restir_reservoir *out) {
    uint32_t W = u->screenSize[0], H = u->screenSize[1];
    v3 cam = ld3(u->cameraPos);
    int32_t pointCount = (int32_t)s->n_points, aliasCount = (int32_t)s->n_alias;
    /* one table over both kinds (point lights first) when it has a column for every light */
    int mixed = pointCount > 0 && s->n_tri_lights > 0 && aliasCount == pointCount + (int32_t)s->n_tri_lights;
    /* lights the table's index space names: the union, else point lights if any, else triangles */
    int32_t nLights = mixed ? pointCount + (int32_t)s->n_tri_lights
                            : (pointCount > 0 ? pointCount : (int32_t)s->n_tri_lights);
    size_t npx = (size_t)W * H;
    int standard = (u->flags & RESTIR_STANDARD_WEIGHTS_FLAG) != 0;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t yy = 0; yy < (int64_t)H; ++yy) {
        uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < W; ++x) {
            texel t = fetch(cur, x, y);
            float albLum = oracle_luminance(t.alb.x, t.alb.y, t.alb.z);
            restir_reservoir res;
            memset(&res, 0, sizeof res);
            uint64_t rng[2];
            pcg_seed(rng, (uint64_t)u->frame, (uint64_t)(uint32_t)(y * 10007u + x));
            if (vdot(t.n, t.n) != 0.0f) {
                for (uint32_t i = 0; i < u->initialLightSampleCount; ++i) {
                    int32_t idx;
                    float prob;
                    int point;
                    if (cand_light) {
                        size_t ci = (size_t)i * npx + (size_t)y * W + x;
                        idx = cand_light[ci];
                        prob = cand_q[ci];
                        if (!(idx >= 0 && idx < nLights && prob > 0.0f)) {
                            const float zero3[3] = {0, 0, 0}, zero4[4] = {0, 0, 0, 0};
                            oracle_add_sample(&res, zero3, zero4, 0.0f, 0, 0.0f, 1.0f, rng);
                            continue;
                        }
                        point = idx < pointCount;
                    } else {
                        float r1 = pcg_float(rng), r2 = pcg_float(rng);
                        /* aliasTableSample, :73-83 */
                        int32_t col = (int32_t)((float)aliasCount * r1);
                        if (col > aliasCount - 1) col = aliasCount - 1;
                        const restir_alias_table_column *c = &s->alias[col];
                        if (c->prob > r2) { idx = col; prob = c->oriProb; }
                        else { idx = c->alias; prob = c->aliasOriProb; }
                        point = mixed ? idx < pointCount : pointCount != 0;
                    }
                    v3 L;
                    float ln4[4] = {0, 0, 0, 0};
                    float lum;
                    int32_t lidx;
                    if (point) {
                        const restir_point_light *pl = &s->points[idx];
                        L = ld3(pl->pos);
                        lum = pl->color_luminance[3];
                        lidx = idx;
                    } else {
                        if (cand_light || mixed) idx -= pointCount;
                        const restir_tri_light *tl = &s->tri_lights[idx];
                        float a = pcg_float(rng), b = pcg_float(rng);
                        float sa = sqrtf(a); /* pickPointOnTriangle, :68-71 */
                        L = vadd(vadd(smul(1.0f - sa, ld3(tl->p1)), smul(sa * (1.0f - b), ld3(tl->p2))),
                                 smul(b * sa, ld3(tl->p3)));
                        lum = tl->emission_luminance[3];
                        lidx = -1 - idx;
                        v3 wi = vnorm(vsub(t.pos, L));
                        v3 nl = ld3(tl->normalArea);
                        if (standard) prob /= tl->normalArea[3]; /* uniform by area */
                        else prob /= fabsf(vdot(wi, nl)) * tl->normalArea[3];
                        ln4[0] = nl.x; ln4[1] = nl.y; ln4[2] = nl.z; ln4[3] = 1.0f;
                    }
                    float pHat = phat_v(t.pos, L, cam, t.n, ld3(ln4), ln4[3] > 0.5f, albLum, lum,
                                        t.rough, t.metal);
                    float Lp[3] = {L.x, L.y, L.z};
                    oracle_add_sample(&res, Lp, ln4, lum, lidx, pHat, prob, rng);
                }
                /* RESTIR_STANDARD_WEIGHTS_FLAG: W = sumW / (M p_hat) after the candidates */
                if (standard && res.samples[0].w > 0.0f)
                    res.samples[0].w = res.samples[0].sumWeights /
                                       ((float)res.numStreamSamples * res.samples[0].pHat);
            }
            /* visibility reuse, :148-160. A reservoir with sumWeights == 0 holds w == 0 and is
             * left unchanged by a shadow hit, so its ray is not traced. */
            if ((u->flags & RESTIR_VISIBILITY_REUSE_FLAG) && res.samples[0].sumWeights != 0.0f) {
                if (test_visibility(s, t.pos, res_pos(&res))) {
                    res.samples[0].w = 0.0f;
                    res.samples[0].sumWeights = 0.0f;
                }
            }
            /* temporal reuse, :163-209 */
            if (u->flags & RESTIR_TEMPORAL_REUSE_FLAG) {
                const float *m = u->prevFrameProjectionViewMatrix;
                float pp[4];
                for (int r = 0; r < 4; ++r)
                    pp[r] = m[0 * 4 + r] * t.pos.x + m[1 * 4 + r] * t.pos.y + m[2 * 4 + r] * t.pos.z + m[3 * 4 + r] * 1.0f;
                pp[0] /= pp[3]; pp[1] /= pp[3]; pp[2] /= pp[3];
                pp[0] = (pp[0] + 1.0f) * 0.5f * (float)W;
                pp[1] = (pp[1] + 1.0f) * 0.5f * (float)H;
                if (pp[0] > 0.0f && pp[1] > 0.0f && pp[0] < (float)W && pp[1] < (float)H) {
                    uint32_t px = (uint32_t)(int32_t)pp[0], py = (uint32_t)(int32_t)pp[1];
                    texel pt = fetch(prev, px, py);
                    v3 pd = vsub(t.pos, pt.pos);
                    if (vdot(pd, pd) < 0.01f) {
                        v3 ad = vsub(t.alb, pt.alb);
                        if (vdot(ad, ad) < 0.01f) {
                            if (vdot(t.n, pt.n) > 0.5f) {
                                restir_reservoir pr = prev_res[(size_t)py * W + px];
                                uint32_t cap = u->temporalSampleCountMultiplier * res.numStreamSamples;
                                if (pr.numStreamSamples > cap) pr.numStreamSamples = cap;
                                float ph = res_phat_at(&pr, t.pos, cam, t.n, albLum, t.rough, t.metal);
                                oracle_combine(&res, &pr, ph, rng);
                            }
                        }
                    }
                }
            }
            out[(size_t)y * W + x] = res;
        }
    }
}

void oracle_pass_a(const oracle_scene *s, const restir_uniforms *u, const restir_gbuffer *cur,
                   const restir_gbuffer *prev, const restir_reservoir *prev_res,
                   restir_reservoir *out) {
    pass_a_impl(s, u, cur, prev, prev_res, NULL, NULL, out);
}

void oracle_pass_a_candidates(const oracle_scene *s, const restir_uniforms *u, const restir_gbuffer *cur,
                              const restir_gbuffer *prev, const restir_reservoir *prev_res,
                              const int32_t *cand_light, const float *cand_q, restir_reservoir *out) {
    pass_a_impl(s, u, cur, prev, prev_res, cand_light, cand_q, out);
}

float oracle_cos_threshold(float degrees) {
    float s, c;
    oracle_sincos(degrees * 0.0174532924f, &s, &c); /* radians(x) = x * pi/180 */
    return c;
}

void oracle_pass_b(const oracle_scene *sc, const restir_uniforms *u, int iter,
                   const restir_gbuffer *g, const restir_reservoir *in, restir_reservoir *out) {
    /* spatialReuse.comp:30-85 */
    (void)sc;
    uint32_t W = u->screenSize[0], H = u->screenSize[1];
    v3 cam = ld3(u->cameraPos);
    float cosThr = oracle_cos_threshold(u->spatialNormalThreshold);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t yy = 0; yy < (int64_t)H; ++yy) {
        uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < W; ++x) {
            texel t = fetch(g, x, y);
            float albLum = oracle_luminance(t.alb.x, t.alb.y, t.alb.z);
            size_t idx = (size_t)y * W + x;
            restir_reservoir res = in[idx];
            uint64_t rng[2];
            pcg_seed(rng, (uint64_t)(uint32_t)(u->frame * 31u + (uint32_t)iter),
                     (uint64_t)(uint32_t)(y * 10007u + x));
            for (uint32_t i = 0; i < u->spatialNeighbors; ++i) {
                float angle = pcg_float(rng) * 2.0f * PI_F;
                float radius = sqrtf(pcg_float(rng)) * u->spatialRadius;
                float sn, cs;
                oracle_sincos(angle, &sn, &cs);
                int32_t ox = (int32_t)floorf(cs * radius), oy = (int32_t)floorf(sn * radius);
                int32_t nx = (int32_t)x + ox, ny = (int32_t)y + oy;
                nx = nx < 0 ? 0 : (nx > (int32_t)W - 1 ? (int32_t)W - 1 : nx);
                ny = ny < 0 ? 0 : (ny > (int32_t)H - 1 ? (int32_t)H - 1 : ny);
                size_t ni = (size_t)ny * W + (size_t)nx;
                float nd = g->depth[ni];
                v3 nn = mk(((const float *)g->normal)[ni], ((const float *)g->normal)[g->plane_stride + ni], ((const float *)g->normal)[2 * g->plane_stride + ni]);
                if (fabsf(nd - t.depth) > u->spatialPosThreshold * fabsf(t.depth) || vdot(nn, t.n) < cosThr)
                    continue;
                restir_reservoir rr = in[ni];
                float ph = res_phat_at(&rr, t.pos, cam, t.n, albLum, t.rough, t.metal);
                oracle_combine(&res, &rr, ph, rng);
            }
            out[idx] = res;
        }
    }
}

void oracle_pass_c(const oracle_scene *s, const restir_uniforms *u, int k, int iter,
                   const restir_gbuffer *g, const restir_reservoir *in, restir_reservoir *out) {
    /* unbiasedReuse.glsl:50-185, NUM_NEIGHBORS = k. Background pixels (zero normal) pass their
     * input through: the reference's result there depends on uninitialised reservoir fields. */
    uint32_t W = u->screenSize[0], H = u->screenSize[1];
    v3 cam = ld3(u->cameraPos);
    int vis = (u->flags & RESTIR_VISIBILITY_REUSE_FLAG) != 0;
    int standard = (u->flags & RESTIR_STANDARD_WEIGHTS_FLAG) != 0;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t yy = 0; yy < (int64_t)H; ++yy) {
        uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < W; ++x) {
            size_t idx = (size_t)y * W + x;
            texel t = fetch(g, x, y);
            restir_reservoir res = in[idx];
            if (vdot(t.n, t.n) == 0.0f) { out[idx] = res; continue; }
            float albLum = oracle_luminance(t.alb.x, t.alb.y, t.alb.z);
            uint64_t rng[2];
            pcg_seed(rng, (uint64_t)(uint32_t)(u->frame * 17u + (uint32_t)iter), (uint64_t)(uint32_t)(y * 10007u + x));
            size_t nidx[16];
            uint32_t nM[16];
            uint32_t origM = res.numStreamSamples;
            for (int i = 0; i < k; ++i) {
                float angle = pcg_float(rng) * 2.0f * PI_F;
                float radius = sqrtf(pcg_float(rng)) * u->spatialRadius;
                float sn, cs;
                oracle_sincos(angle, &sn, &cs);
                int32_t ox = (int32_t)roundf(cs * radius), oy = (int32_t)roundf(sn * radius);
                int32_t nx = (int32_t)x + ox, ny = (int32_t)y + oy;
                nx = nx < 0 ? 0 : (nx > (int32_t)W - 1 ? (int32_t)W - 1 : nx);
                ny = ny < 0 ? 0 : (ny > (int32_t)H - 1 ? (int32_t)H - 1 : ny);
                size_t ni = (size_t)ny * W + (size_t)nx;
                restir_reservoir rr = in[ni];
                nidx[i] = ni;
                nM[i] = rr.numStreamSamples;
                res.numStreamSamples += rr.numStreamSamples;
                float ph = res_phat_at(&rr, t.pos, cam, t.n, albLum, t.rough, t.metal);
                float weight = ph * rr.samples[0].w * (float)rr.numStreamSamples;
                if (weight > 0.0f)
                    update_at(&res, weight, res_pos(&rr), rr.samples[0].normal,
                              rr.samples[0].position_emissionLum[3], rr.samples[0].lightIndex, ph,
                              rr.samples[0].w, rng);
            }
            v3 lp = res_pos(&res);
            uint32_t numSamples = origM;
            for (int j = 0; j < k; ++j) {
                size_t ni = nidx[j], P = g->plane_stride;
                v3 nwp = mk(g->position[ni], g->position[P + ni], g->position[2 * P + ni]);
                v3 nn = mk(((const float *)g->normal)[ni], ((const float *)g->normal)[P + ni], ((const float *)g->normal)[2 * P + ni]);
                if (vdot(vsub(lp, nwp), nn) < 0.0f) continue;
                if (vis && test_visibility(s, nwp, lp)) continue;
                numSamples += nM[j];
            }
            if (vis && test_visibility(s, t.pos, lp)) numSamples = 0;
            if (numSamples > 0 && (!standard || res.samples[0].pHat > 0.0f)) {
                res.samples[0].w = res.samples[0].sumWeights / ((float)numSamples * res.samples[0].pHat);
            } else {
                res.samples[0].w = 0.0f;
                res.samples[0].sumWeights = 0.0f;
            }
            /* RESTIR_STANDARD_WEIGHTS_FLAG: keep sumW = p_hat W M, the weight a later pass gives
             * every other input reservoir (the reference keeps p_hat W Z) */
            if (standard)
                res.samples[0].sumWeights = res.samples[0].w * (float)res.numStreamSamples * res.samples[0].pHat;
            out[idx] = res;
        }
    }
}

static void shade_impl(const oracle_scene *s, const float camf[3], float gamma, int visibility,
                       const restir_gbuffer *g, const restir_reservoir *res, float *rgb, uint64_t rs) {
    /* lighting.frag:49-71, 103; with `visibility`, the sample of a pixel whose weight is nonzero
     * is shadow-tested from the pixel first (testVisibility) and a shadowed one shades to 0 */
    uint32_t W = g->width, H = g->height;
    v3 cam = ld3(camf);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        texel t = fetch(g, (uint32_t)(i % W), (uint32_t)(i / W));
        float c[3] = {0, 0, 0};
        if (vdot(t.n, t.n) != 0.0f) {
            const restir_light_sample *sm = &res[i].samples[0];
            static const float zero3[3] = {0.0f, 0.0f, 0.0f};
            const float *em = zero3; /* an index outside the tables shades black */
            if (sm->lightIndex < 0) {
                if ((uint32_t)(-1 - sm->lightIndex) < s->n_tri_lights)
                    em = s->tri_lights[-1 - sm->lightIndex].emission_luminance;
            } else if ((uint32_t)sm->lightIndex < s->n_points) {
                em = s->points[sm->lightIndex].color_luminance;
            }
            float alb[3] = {t.alb.x, t.alb.y, t.alb.z};
            phat_full_v(t.pos, ld3(sm->position_emissionLum), cam, t.n, ld3(sm->normal),
                        sm->normal[3] > 0.5f, alb, em, t.rough, t.metal, c);
            for (int j = 0; j < 3; ++j) c[j] = c[j] * sm->w;
            if (visibility && sm->w != 0.0f && test_visibility(s, t.pos, ld3(sm->position_emissionLum)))
                c[0] = c[1] = c[2] = 0.0f;
        }
        if (t.emissive > 0.5f) { c[0] = t.alb.x; c[1] = t.alb.y; c[2] = t.alb.z; }
        if (gamma != 1.0f)
            for (int j = 0; j < 3; ++j) c[j] = powf(c[j], 1.0f / gamma);
        rgb[i] = c[0]; rgb[rs + i] = c[1]; rgb[2 * rs + i] = c[2];
    }
}
void oracle_shade(const oracle_scene *s, const float camf[3], float gamma, const restir_gbuffer *g,
                  const restir_reservoir *res, float *rgb, uint64_t rs) {
    shade_impl(s, camf, gamma, 0, g, res, rgb, rs);
}
void oracle_shade_visibility(const oracle_scene *s, const float camf[3], float gamma, const restir_gbuffer *g,
                             const restir_reservoir *res, float *rgb, uint64_t rs) {
    shade_impl(s, camf, gamma, 1, g, res, rgb, rs);
}

/* ------------------------------------------------------------------ primary G-buffer ------ */
typedef struct { v3 pos, fwd, right, down; float sx, sy, A, B; } cam_basis;
static cam_basis make_basis(const restir_camera *c, float aspect) {
    /* camera.h:25-50 */
    cam_basis b;
    b.pos = ld3(c->position);
    v3 fwd = vnorm(vsub(ld3(c->lookAt), b.pos));
    v3 right = vnorm(vcross(fwd, ld3(c->worldUp)));
    v3 up = vcross(right, fwd);
    b.fwd = fwd; b.right = right; b.down = mk(-up.x, -up.y, -up.z);
    float f = 1.0f / tanf(0.5f * c->fovYRadians);
    b.sx = aspect / f;
    b.sy = 1.0f / f;
    b.A = -c->zFar / (c->zNear - c->zFar);
    b.B = c->zNear * c->zFar / (c->zNear - c->zFar);
    return b;
}
void oracle_projection_view(const restir_camera *c, float aspect, float out[16]) {
    cam_basis b = make_basis(c, aspect);
    float f = 1.0f / tanf(0.5f * c->fovYRadians);
    float view[4][4] = {{b.right.x, b.right.y, b.right.z, -vdot(b.right, b.pos)},
                        {b.down.x, b.down.y, b.down.z, -vdot(b.down, b.pos)},
                        {b.fwd.x, b.fwd.y, b.fwd.z, -vdot(b.fwd, b.pos)},
                        {0, 0, 0, 1}};
    float proj[4][4] = {{f / aspect, 0, 0, 0}, {0, f, 0, 0}, {0, 0, b.A, b.B}, {0, 0, 1, 0}};
    for (int r = 0; r < 4; ++r)
        for (int cc = 0; cc < 4; ++cc) {
            float acc = 0.0f;
            for (int k = 0; k < 4; ++k) acc += proj[r][k] * view[k][cc];
            out[cc * 4 + r] = acc;
        }
}
static inline int ray_tri_closest(const restir_triangle *t, v3 o, v3 d, float *tout) {
    v3 p1 = ld3(t->p1);
    v3 e1 = vsub(ld3(t->p2), p1);
    v3 e2 = vsub(ld3(t->p3), p1);
    v3 p = vcross(d, e2);
    float f = 1.0f / vdot(e1, p);
    v3 s = vsub(o, p1);
    float bx = f * vdot(s, p);
    if (bx < 0.0f || bx > 1.0f) return 0;
    v3 q = vcross(s, e1);
    float by = f * vdot(d, q);
    if (by < 0.0f || by + bx > 1.0f) return 0;
    float tt = f * vdot(e2, q);
    if (!(tt > 0.0f)) return 0;
    *tout = tt;
    return 1;
}
void oracle_primary_gbuffer(const restir_triangle *tris, uint32_t n_tris, const restir_camera *cam,
                            const int32_t *tri_material, const float *mats, float *position,
                            float *normal, float *albedo, float *material, float *depth,
                            uint64_t P, uint32_t W, uint32_t H) {
    cam_basis b = make_basis(cam, (float)W / (float)H);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t yy = 0; yy < (int64_t)H; ++yy) {
        for (uint32_t x = 0; x < W; ++x) {
            size_t i = (size_t)yy * W + x;
            float ndx = ((float)x + 0.5f) / (float)W * 2.0f - 1.0f;
            float ndy = ((float)yy + 0.5f) / (float)H * 2.0f - 1.0f;
            v3 d = vnorm(vadd(vadd(b.fwd, vmul(b.right, ndx * b.sx)), vmul(b.down, ndy * b.sy)));
            float best = INFINITY;
            int32_t bi = -1;
            for (uint32_t t = 0; t < n_tris; ++t) {
                float tt;
                if (ray_tri_closest(&tris[t], b.pos, d, &tt) && tt < best) { best = tt; bi = (int32_t)t; }
            }
            if (bi < 0) {
                for (int c = 0; c < 3; ++c) { position[c * P + i] = 0; normal[c * P + i] = 0; }
                for (int c = 0; c < 4; ++c) albedo[c * P + i] = 0;
                material[i] = 0; material[P + i] = 0;
                depth[i] = 1.0f;
                continue;
            }
            v3 pos = vadd(b.pos, vmul(d, best));
            const restir_triangle *tr = &tris[bi];
            v3 p1 = ld3(tr->p1);
            v3 n = vnorm(vcross(vsub(ld3(tr->p2), p1), vsub(ld3(tr->p3), p1)));
            if (vdot(n, d) > 0.0f) n = mk(-n.x, -n.y, -n.z);
            const float *m = mats + 8 * (size_t)tri_material[bi];
            position[i] = pos.x; position[P + i] = pos.y; position[2 * P + i] = pos.z;
            normal[i] = n.x; normal[P + i] = n.y; normal[2 * P + i] = n.z;
            albedo[i] = m[0]; albedo[P + i] = m[1]; albedo[2 * P + i] = m[2]; albedo[3 * P + i] = m[3];
            material[i] = m[4]; material[P + i] = m[5];
            float z = best * vdot(d, b.fwd);
            depth[i] = (b.A * z + b.B) / z;
        }
    }
}

/* --------------------------------------------------------------- FP64 ground truth ------- */
static int occluded_f64_brute(const oracle_scene *s, const double a[3], const double bpt[3]) {
    double d[3], o[3];
    for (int k = 0; k < 3; ++k) d[k] = bpt[k] - a[k];
    double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) {
        double off = 0.001 * d[k] / len;
        o[k] = a[k] + off;
        d[k] = d[k] - 2.0 * off;
    }
    for (uint32_t t = 0; t < s->n_tris; ++t)
        if (seg_tri_f64(&s->tris[t], o, d, NULL)) return 1;
    return 0;
}
static int box_f64(const double o[3], const double inv[3], const float *mn, const float *mx) {
    double t0 = -INFINITY, t1 = INFINITY;
    for (int k = 0; k < 3; ++k) {
        double a = ((double)mn[k] - o[k]) * inv[k], b = ((double)mx[k] - o[k]) * inv[k];
        if (a != a || b != b) { /* direction component 0 and origin on the slab plane */
            if (o[k] < mn[k] || o[k] > mx[k]) return 0;
            continue;
        }
        double lo = a < b ? a : b, hi = a < b ? b : a;
        if (lo > t0) t0 = lo;
        if (hi < t1) t1 = hi;
    }
    /* conservative slack so FP64 box rounding never culls a true hit */
    return t0 <= t1 + 1e-9 && t0 < 1.0 + 1e-9 && t1 > -1e-9;
}
static int occluded_f64_bvh(const oracle_scene *s, const double a[3], const double bpt[3]) {
    double d[3], o[3], inv[3];
    for (int k = 0; k < 3; ++k) d[k] = bpt[k] - a[k];
    double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) {
        double off = 0.001 * d[k] / len;
        o[k] = a[k] + off;
        d[k] = d[k] - 2.0 * off;
        inv[k] = 1.0 / d[k];
    }
    int32_t stack[ORACLE_STACK];
    int top = 1;
    stack[0] = 0;
    while (top > 0) {
        const restir_aabb_tree_node *nd = &s->nodes[stack[--top]];
        int32_t ch[2] = {nd->leftChild, nd->rightChild};
        const float *mn[2] = {nd->leftAabbMin, nd->rightAabbMin};
        const float *mx[2] = {nd->leftAabbMax, nd->rightAabbMax};
        for (int c = 0; c < 2; ++c) {
            if (!box_f64(o, inv, mn[c], mx[c])) continue;
            if (ch[c] < 0) {
                if (seg_tri_f64(&s->tris[~ch[c]], o, d, NULL)) return 1;
            } else if (top < ORACLE_STACK) {
                stack[top++] = ch[c];
            }
        }
    }
    return 0;
}
static void brdf_color_f64(const double P[3], const double L[3], const double C[3],
                           const double N[3], const double *LN, const double alb[3],
                           double rough, double metal, double out[3], double *geom) {
    double wi[3] = {L[0] - P[0], L[1] - P[1], L[2] - P[2]};
    out[0] = out[1] = out[2] = 0.0;
    *geom = 0.0;
    if (wi[0] * N[0] + wi[1] * N[1] + wi[2] * N[2] < 0.0) return;
    double d2 = wi[0] * wi[0] + wi[1] * wi[1] + wi[2] * wi[2], dl = sqrt(d2);
    for (int k = 0; k < 3; ++k) wi[k] /= dl;
    double wo[3] = {C[0] - P[0], C[1] - P[1], C[2] - P[2]};
    double lo = sqrt(wo[0] * wo[0] + wo[1] * wo[1] + wo[2] * wo[2]);
    for (int k = 0; k < 3; ++k) wo[k] /= lo;
    double h[3] = {wi[0] + wo[0], wi[1] + wo[1], wi[2] + wo[2]};
    double lh = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    for (int k = 0; k < 3; ++k) h[k] /= lh;
    double ci = N[0] * wi[0] + N[1] * wi[1] + N[2] * wi[2];
    double co = N[0] * wo[0] + N[1] * wo[1] + N[2] * wo[2];
    double ch = N[0] * h[0] + N[1] * h[1] + N[2] * h[2];
    double cih = wi[0] * h[0] + wi[1] * h[1] + wi[2] * h[2];
    if (ci < 0.0) return;
    double df, fih, gsds;
    brdf_parts_d(ci, co, ch, cih, rough, metal, &df, &fih, &gsds);
    for (int k = 0; k < 3; ++k) {
        double sc = 0.04 * (1.0 - metal) + alb[k] * metal;
        double fs = sc * (1.0 - fih) + fih;
        out[k] = alb[k] * df + fs * gsds;
    }
    double g = ci / d2;
    if (LN) g *= fabs(wi[0] * LN[0] + wi[1] * LN[1] + wi[2] * LN[2]);
    *geom = g;
}
void oracle_ground_truth_f64(const oracle_scene *s, const float camf[3], const restir_gbuffer *g,
                             int use_bvh, uint32_t n_tri_samples, const uint8_t *mask,
                             double *rgb, uint64_t rs) {
    uint32_t W = g->width, H = g->height;
    double C[3] = {camf[0], camf[1], camf[2]};
    uint32_t sq = 1;
    while (sq * sq < n_tri_samples) ++sq;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        if (mask && !mask[i]) continue;
        texel t = fetch(g, (uint32_t)(i % W), (uint32_t)(i / W));
        double acc[3] = {0, 0, 0};
        if (t.emissive > 0.5f) {
            acc[0] = t.alb.x; acc[1] = t.alb.y; acc[2] = t.alb.z;
        } else if (vdot(t.n, t.n) != 0.0f) {
            double P[3] = {t.pos.x, t.pos.y, t.pos.z}, N[3] = {t.n.x, t.n.y, t.n.z};
            double alb[3] = {t.alb.x, t.alb.y, t.alb.z};
            for (uint32_t l = 0; l < s->n_points; ++l) {
                const restir_point_light *pl = &s->points[l];
                double L[3] = {pl->pos[0], pl->pos[1], pl->pos[2]};
                double f[3], geom;
                brdf_color_f64(P, L, C, N, NULL, alb, t.rough, t.metal, f, &geom);
                if (geom <= 0.0) continue;
                if (use_bvh ? occluded_f64_bvh(s, P, L) : occluded_f64_brute(s, P, L)) continue;
                for (int k = 0; k < 3; ++k) acc[k] += pl->color_luminance[k] * f[k] * geom;
            }
            /* triangle lights: stratified sqrt-warp samples of the area measure */
            for (uint32_t l = 0; l < s->n_tri_lights && n_tri_samples > 0; ++l) {
                const restir_tri_light *tl = &s->tri_lights[l];
                double LN[3] = {tl->normalArea[0], tl->normalArea[1], tl->normalArea[2]};
                double area = tl->normalArea[3];
                double sum[3] = {0, 0, 0};
                for (uint32_t a = 0; a < sq; ++a)
                    for (uint32_t bb = 0; bb < sq; ++bb) {
                        double r1 = (a + 0.5) / sq, r2 = (bb + 0.5) / sq, sr = sqrt(r1);
                        double L[3];
                        for (int k = 0; k < 3; ++k)
                            L[k] = (1.0 - sr) * tl->p1[k] + sr * (1.0 - r2) * tl->p2[k] + r2 * sr * tl->p3[k];
                        double f[3], geom;
                        brdf_color_f64(P, L, C, N, LN, alb, t.rough, t.metal, f, &geom);
                        if (geom <= 0.0) continue;
                        if (use_bvh ? occluded_f64_bvh(s, P, L) : occluded_f64_brute(s, P, L)) continue;
                        for (int k = 0; k < 3; ++k) sum[k] += tl->emission_luminance[k] * f[k] * geom;
                    }
                for (int k = 0; k < 3; ++k) acc[k] += sum[k] * area / ((double)sq * sq);
            }
        }
        rgb[i] = acc[0]; rgb[rs + i] = acc[1]; rgb[2 * rs + i] = acc[2];
    }
}

/* ------------------------------------------------------------- light clusters (k-means) ---- */
/* Seeded k-means over point-light positions and per-cluster Vose tables over luminance.
 * Initial centroids: partial Fisher-Yates with PCG32(seed, seq 0x6b6d65616e73); `iterations`
 * Lloyd steps (squared distance in float, ties to the lower centroid index; means accumulated
 * in double in light order; empty clusters keep their centroid), then a final assignment.
 * Lights are listed by cluster in index order; alias entries are local to their cluster. */
static void vose_local(const float *w, uint32_t n, restir_alias_table_column *out) {
    /* misc.cpp:418-497 on an arbitrary weight list */
    float *p = (float *)malloc(sizeof(float) * n);
    uint32_t *big = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    uint32_t *small = (uint32_t *)malloc(sizeof(uint32_t) * (n + 1));
    size_t bh = 0, bt = 0, sh = 0, st = 0, cap = n + 1;
    float total = 0.0f;
    for (uint32_t i = 0; i < n; ++i) { total += w[i]; p[i] = w[i]; }
    for (uint32_t i = 0; i < n; ++i) {
        out[i].prob = 0.0f;
        out[i].alias = -1;
        out[i].oriProb = p[i] / total;
        out[i].aliasOriProb = 0.0f;
        p[i] = (float)n * p[i] / total;
        if (p[i] >= 1.0f) big[bt++ % cap] = i; else small[st++ % cap] = i;
    }
    while (bh != bt && sh != st) {
        uint32_t g = big[bh++ % cap], l = small[sh++ % cap];
        out[l].prob = p[l];
        out[l].alias = (int32_t)g;
        p[g] = (p[g] + p[l]) - 1.0f;
        if (p[g] < 1.0f) small[st++ % cap] = g; else big[bt++ % cap] = g;
    }
    while (bh != bt) { uint32_t g = big[bh++ % cap]; out[g].prob = 1.0f; out[g].alias = (int32_t)g; }
    while (sh != st) { uint32_t l = small[sh++ % cap]; out[l].prob = 1.0f; out[l].alias = (int32_t)l; }
    for (uint32_t i = 0; i < n; ++i) out[i].aliasOriProb = out[out[i].alias].oriProb;
    free(p); free(big); free(small);
}

int oracle_build_light_clusters(const restir_point_light *pts, uint32_t n, uint32_t k, uint64_t seed,
                                uint32_t iterations, int32_t *light_cluster, float *light_pwithin,
                                int32_t *cl_begin, int32_t *cl_count, int32_t *cl_lights,
                                restir_alias_table_column *cl_alias, float *centroids) {
    if (k == 0 || k > n) return -1;
    uint32_t *perm = (uint32_t *)malloc(sizeof(uint32_t) * n);
    float *c = (float *)malloc(sizeof(float) * 3 * k);
    double *sum = (double *)malloc(sizeof(double) * 3 * k);
    uint32_t *cnt = (uint32_t *)malloc(sizeof(uint32_t) * k);
    uint32_t *asg = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    uint64_t rng[2];
    pcg_seed(rng, seed, 0x6b6d65616e73ULL);
    for (uint32_t i = 0; i < k; ++i) {
        uint32_t j = i + pcg_next(rng) % (n - i);
        uint32_t t = perm[i]; perm[i] = perm[j]; perm[j] = t;
        for (int d = 0; d < 3; ++d) c[3 * i + d] = pts[perm[i]].pos[d];
    }
    for (uint32_t it = 0;; ++it) {
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < (int64_t)n; ++l) {
            float x = pts[l].pos[0], y = pts[l].pos[1], z = pts[l].pos[2];
            uint32_t best = 0;
            float bd = 0.0f;
            for (uint32_t j = 0; j < k; ++j) {
                float dx = x - c[3 * j], dy = y - c[3 * j + 1], dz = z - c[3 * j + 2];
                float d = (dx * dx + dy * dy) + dz * dz;
                if (j == 0 || d < bd) { bd = d; best = j; }
            }
            asg[l] = best;
        }
        if (it == iterations) break;
        memset(sum, 0, sizeof(double) * 3 * k);
        memset(cnt, 0, sizeof(uint32_t) * k);
        for (uint32_t l = 0; l < n; ++l) {
            for (int d = 0; d < 3; ++d) sum[3 * asg[l] + d] += (double)pts[l].pos[d];
            cnt[asg[l]]++;
        }
        for (uint32_t j = 0; j < k; ++j)
            if (cnt[j]) for (int d = 0; d < 3; ++d) c[3 * j + d] = (float)(sum[3 * j + d] / cnt[j]);
    }
    memset(cnt, 0, sizeof(uint32_t) * k);
    for (uint32_t l = 0; l < n; ++l) cnt[asg[l]]++;
    int32_t off = 0;
    for (uint32_t j = 0; j < k; ++j) { cl_begin[j] = off; cl_count[j] = (int32_t)cnt[j]; off += (int32_t)cnt[j]; }
    memset(cnt, 0, sizeof(uint32_t) * k);
    for (uint32_t l = 0; l < n; ++l) {
        light_cluster[l] = (int32_t)asg[l];
        cl_lights[cl_begin[asg[l]] + (int32_t)cnt[asg[l]]++] = (int32_t)l;
    }
    float *w = (float *)malloc(sizeof(float) * n);
    for (uint32_t j = 0; j < k; ++j) {
        int32_t b = cl_begin[j], m = cl_count[j];
        if (m == 0) continue;
        for (int32_t i = 0; i < m; ++i) w[i] = pts[cl_lights[b + i]].color_luminance[3];
        vose_local(w, (uint32_t)m, cl_alias + b);
        for (int32_t i = 0; i < m; ++i) light_pwithin[cl_lights[b + i]] = cl_alias[b + i].oriProb;
    }
    if (centroids) memcpy(centroids, c, sizeof(float) * 3 * k);
    free(w); free(perm); free(c); free(sum); free(cnt); free(asg);
    return 0;
}

/* ------------------------------------------------------------------ learned proposal ---- */
/* MLP light proposal (no reference counterpart). Features (114 used of 128): the surface position
 * normalised to the scene box, the normal and the unit view direction; raw position and normal
 * (6), then for each of the 9 scalars s and k = 0..5: sin(2^k*pi*s), cos(2^k*pi*s).
 * Network: 128 -> 128 -> 128 -> 128 -> K with ReLU; weights stored transposed W[out][in].
 * With emulate_bf16 the forward/backward round at the points the GPU stores bf16 tensors:
 * hidden activations after ReLU, and dlogits / dhidden before each GEMM that consumes them. */
static inline float bf16_round(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x; /* inf / nan unchanged */
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

void oracle_mlp_features(const restir_gbuffer *g, const float cam[3], const float lo[3],
                         const float hi[3], const uint32_t *pixels, uint32_t n, float *out) {
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < (int64_t)n; ++b) {
        texel t = fetch(g, pixels[b] % g->width, pixels[b] / g->width);
        double s[9];
        double p[3] = {t.pos.x, t.pos.y, t.pos.z};
        for (int k = 0; k < 3; ++k) {
            double c = 0.5 * ((double)lo[k] + hi[k]), h = 0.5 * ((double)hi[k] - lo[k]);
            s[k] = (p[k] - c) / h;
        }
        s[3] = t.n.x; s[4] = t.n.y; s[5] = t.n.z;
        double v[3] = {cam[0] - p[0], cam[1] - p[1], cam[2] - p[2]};
        double vl = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; ++k) s[6 + k] = vl > 0 ? v[k] / vl : 0.0;
        float *o = out + 128 * b;
        for (int k = 0; k < 6; ++k) o[k] = (float)s[k];
        for (int i = 0; i < 9; ++i)
            for (int f = 0; f < 6; ++f) {
                double a = M_PI * (double)(1 << f) * s[i];
                o[6 + 12 * i + 2 * f] = (float)sin(a);
                o[6 + 12 * i + 2 * f + 1] = (float)cos(a);
            }
        for (int k = 114; k < 128; ++k) o[k] = 0.0f;
    }
}

/* offsets of the four weight matrices / bias vectors in the flat parameter arrays */
static void mlp_offsets(uint32_t K, size_t wo[4], size_t bo[4]) {
    wo[0] = 0; wo[1] = 128 * 128; wo[2] = 2 * 128 * 128; wo[3] = 3 * 128 * 128;
    bo[0] = 0; bo[1] = 128; bo[2] = 256; bo[3] = 384;
    (void)K;
}

static void mlp_forward_one(const float *W, const float *B, uint32_t K, const float *x, int emu,
                            double h[3][128], double *z) {
    size_t wo[4], bo[4];
    mlp_offsets(K, wo, bo);
    const double *prev = NULL;
    double in0[128];
    for (int k = 0; k < 128; ++k) in0[k] = x[k];
    prev = in0;
    for (int l = 0; l < 3; ++l) {
        for (int o = 0; o < 128; ++o) {
            double acc = B[bo[l] + o];
            const float *w = W + wo[l] + 128 * (size_t)o;
            for (int k = 0; k < 128; ++k) acc += (double)w[k] * prev[k];
            double r = acc > 0.0 ? acc : 0.0;
            h[l][o] = emu ? (double)bf16_round((float)r) : r;
        }
        prev = h[l];
    }
    for (uint32_t o = 0; o < K; ++o) {
        double acc = B[bo[3] + o];
        const float *w = W + wo[3] + 128 * (size_t)o;
        for (int k = 0; k < 128; ++k) acc += (double)w[k] * prev[k];
        z[o] = acc;
    }
}

void oracle_mlp_forward(const float *W, const float *B, uint32_t K, const float *X, uint32_t n,
                        int emulate_bf16, double *logits) {
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < (int64_t)n; ++b) {
        double h[3][128];
        mlp_forward_one(W, B, K, X + 128 * b, emulate_bf16, h, logits + (size_t)K * b);
    }
}

/* Mean cross-entropy over samples with mask != 0 against target clusters, and its gradient. */
double oracle_mlp_loss_grad(const float *W, const float *B, uint32_t K, const float *X,
                            const int32_t *target, const uint8_t *mask, uint32_t n,
                            int emulate_bf16, double *gW, double *gB) {
    size_t wo[4], bo[4];
    mlp_offsets(K, wo, bo);
    size_t nW = 3 * 128 * 128 + (size_t)K * 128, nB = 384 + K;
    memset(gW, 0, sizeof(double) * nW);
    memset(gB, 0, sizeof(double) * nB);
    uint32_t nv = 0;
    for (uint32_t b = 0; b < n; ++b) nv += mask[b] ? 1u : 0u;
    if (nv == 0) return 0.0;
    double loss = 0.0;
    double *z = (double *)malloc(sizeof(double) * K), *dz = (double *)malloc(sizeof(double) * K);
    for (uint32_t b = 0; b < n; ++b) {
        if (!mask[b]) continue;
        double h[3][128], dh[128], dprev[128];
        const float *x = X + 128 * (size_t)b;
        mlp_forward_one(W, B, K, x, emulate_bf16, h, z);
        double mx = z[0];
        for (uint32_t o = 1; o < K; ++o) mx = z[o] > mx ? z[o] : mx;
        double se = 0.0;
        for (uint32_t o = 0; o < K; ++o) se += exp(z[o] - mx);
        loss += -(z[target[b]] - mx - log(se));
        for (uint32_t o = 0; o < K; ++o) {
            double g = (exp(z[o] - mx) / se - (o == (uint32_t)target[b] ? 1.0 : 0.0)) / nv;
            dz[o] = emulate_bf16 ? (double)bf16_round((float)g) : g;
        }
        /* layer 4 */
        for (uint32_t o = 0; o < K; ++o) {
            gB[bo[3] + o] += dz[o];
            for (int k = 0; k < 128; ++k) gW[wo[3] + 128 * (size_t)o + k] += dz[o] * h[2][k];
        }
        for (int k = 0; k < 128; ++k) {
            double acc = 0.0;
            for (uint32_t o = 0; o < K; ++o) acc += dz[o] * W[wo[3] + 128 * (size_t)o + k];
            acc = h[2][k] > 0.0 ? acc : 0.0;
            dh[k] = emulate_bf16 ? (double)bf16_round((float)acc) : acc;
        }
        for (int l = 2; l >= 0; --l) {
            const double *in = NULL;
            double in0[128];
            if (l == 0) { for (int k = 0; k < 128; ++k) in0[k] = x[k]; in = in0; }
            else in = h[l - 1];
            for (int o = 0; o < 128; ++o) {
                gB[bo[l] + o] += dh[o];
                for (int k = 0; k < 128; ++k) gW[wo[l] + 128 * (size_t)o + k] += dh[o] * in[k];
            }
            if (l == 0) break;
            for (int k = 0; k < 128; ++k) {
                double acc = 0.0;
                for (int o = 0; o < 128; ++o) acc += dh[o] * W[wo[l] + 128 * (size_t)o + k];
                acc = h[l - 1][k] > 0.0 ? acc : 0.0;
                dprev[k] = emulate_bf16 ? (double)bf16_round((float)acc) : acc;
            }
            memcpy(dh, dprev, sizeof dh);
        }
    }
    free(z); free(dz);
    return loss / nv;
}

/* Adam (Kingma & Ba) with bias correction at step t >= 1. */
void oracle_adam(double *p, const double *g, double *m, double *v, size_t n, double lr, double b1,
                 double b2, double eps, uint32_t t) {
    double c1 = 1.0 - pow(b1, (double)t), c2 = 1.0 - pow(b2, (double)t);
    for (size_t i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        p[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
    }
}
