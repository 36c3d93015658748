// truth.hip — FP64 ground truth of the direct lighting at listed pixels: sum over every light of
// Le * BRDF * G * V, the integral that ReSTIR estimates (restirUtils.glsl:30-55 integrand,
// disneyBRDF.glsl:5-90 BRDF, visibilityTest.glsl:1-30 segment convention). It is the reference
// of the convergence/error metric, so everything (shading, traversal, intersection) runs in
// double precision; triangle lights are integrated with sq x sq stratified samples of the
// sqrt-warped barycentric map (restirOmni.glsl:68-71) per light.
//
// Work split: workgroup (x, y) covers 256 listed pixels and light chunk y; each lane writes its
// chunk's partial sums, and k_truth_reduce adds the chunks in order (deterministic).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace rgt {

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ d3 ld3(const float *p) { return d3{p[0], p[1], p[2]}; }
__device__ __forceinline__ d3 ld3(const float4 &v) { return d3{v.x, v.y, v.z}; }

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ double schlick(double c) {  // (1 - clamp(c, 0, 1))^5
    const double m = c < 0.0 ? 1.0 : (c > 1.0 ? 0.0 : 1.0 - c);
    return m * m * m * m * m;
}
__device__ __forceinline__ double smith_ggx(double c, double a2) {
    return 1.0 / (fabs(c) + fmax(sqrt(a2 + c * c - a2 * c * c), 0.0001));
}

// Le * BRDF(rgb) * cosIn / d^2 (* |cos at the light| for area lights) at surface point P.
__device__ void integrand(d3 P, d3 N, d3 C, d3 L, const d3 *LN, d3 alb, double rough, double metal,
                          d3 Le, d3 &out) {
    out = d3{0.0, 0.0, 0.0};
    d3 wi = L - P;
    if (dot(wi, N) < 0.0) return;
    const double d2 = dot(wi, wi), dl = sqrt(d2);
    wi = d3{wi.x / dl, wi.y / dl, wi.z / dl};
    d3 wo = C - P;
    const double lo = sqrt(dot(wo, wo));
    wo = d3{wo.x / lo, wo.y / lo, wo.z / lo};
    d3 h = wi + wo;
    const double lh = sqrt(dot(h, h));
    h = d3{h.x / lh, h.y / lh, h.z / lh};
    const double ci = dot(N, wi), co = dot(N, wo), ch = dot(N, h), cih = dot(wi, h);
    if (ci < 0.0) return;
    const double fd90 = 0.5 + 2.0 * cih * cih * rough;
    const double fin = schlick(ci), fout = schlick(co);
    const double diff = ((1.0 - fin) + fd90 * fin) * ((1.0 - fout) + fd90 * fout) * (1.0 - metal) / kPi;
    const double fih = schlick(cih);
    double a = rough * rough;
    a = a < 0.001 ? 0.001 : a;
    const double a2 = a * a;
    const double t = 1.0 + (a2 - 1.0) * ch * ch;
    const double gsds = smith_ggx(ci, a2) * smith_ggx(co, a2) * (a2 / (kPi * t * t));
    double g = ci / d2;
    if (LN) g *= fabs(dot(wi, *LN));
    const double al[3] = {alb.x, alb.y, alb.z}, le[3] = {Le.x, Le.y, Le.z};
    double o[3];
    for (int k = 0; k < 3; ++k) {
        const double sc = 0.04 * (1.0 - metal) + al[k] * metal;
        const double fs = sc * (1.0 - fih) + fih;
        o[k] = le[k] * (al[k] * diff + fs * gsds) * g;
    }
    out = d3{o[0], o[1], o[2]};
}

// slab test of the segment o + t d, t in [0, 1], with a 1e-9 slack so rounding never culls a
// crossing; a zero direction component with the origin on a slab plane tests the origin
__device__ __forceinline__ bool box(d3 o, d3 inv, const float4 &mn, const float4 &mx) {
    double t0 = -INFINITY, t1 = INFINITY;
    const double oo[3] = {o.x, o.y, o.z}, iv[3] = {inv.x, inv.y, inv.z};
    const double lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    for (int k = 0; k < 3; ++k) {
        const double a = (lo[k] - oo[k]) * iv[k], b = (hi[k] - oo[k]) * iv[k];
        if (a != a || b != b) {
            if (oo[k] < lo[k] || oo[k] > hi[k]) return false;
            continue;
        }
        const double l = a < b ? a : b, h = a < b ? b : a;
        if (l > t0) t0 = l;
        if (h < t1) t1 = h;
    }
    return t0 <= t1 + 1e-9 && t0 < 1.0 + 1e-9 && t1 > -1e-9;
}

// Moller-Trumbore on the segment: barycentrics in [0, 1] (closed), t in (0, 1) (open)
__device__ __forceinline__ bool seg_tri(const float4 *tri, d3 o, d3 d) {
    const d3 p1 = ld3(tri[0]), e1 = ld3(tri[1]) - p1, e2 = ld3(tri[2]) - p1, s = o - p1;
    const d3 p = cross(d, e2);
    const double det = dot(e1, p);
    if (det == 0.0) return false;
    const double f = 1.0 / det;
    const double u = f * dot(s, p);
    const d3 q = cross(s, e1);
    const double v = f * dot(d, q);
    const double t = f * dot(e2, q);
    const double w = 1.0 - u - v;
    return u >= 0.0 && v >= 0.0 && w >= 0.0 && t > 0.0 && t < 1.0;
}

// any triangle crossing the segment a -> b shrunk by 0.001 at both ends
template <int STACK>
__device__ bool occluded(const rk::SceneView &sc, d3 a, d3 b, int32_t *stk) {
    d3 d = b - a;
    const double len = sqrt(dot(d, d));
    const d3 off = d3{0.001 * d.x / len, 0.001 * d.y / len, 0.001 * d.z / len};
    const d3 o = a + off;
    d = d - off * 2.0;
    const d3 inv = d3{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
    int top = 1;
    stk[0] = 0;
    while (top > 0) {
        const rt::Node &nd = sc.nodes[stk[(--top) * 256]];
        const int32_t ch[2] = {nd.left, nd.right};
        const float4 mn[2] = {nd.lmin, nd.rmin}, mx[2] = {nd.lmax, nd.rmax};
        for (int c = 0; c < 2; ++c) {
            if (!box(o, inv, mn[c], mx[c])) continue;
            if (ch[c] < 0) {
                if (seg_tri(sc.tris + 3 * static_cast<size_t>(~ch[c]), o, d)) return true;
            } else if (top < STACK) {
                stk[(top++) * 256] = ch[c];
            }
        }
    }
    return false;
}

struct TruthArgs {
    rk::SceneView sc;
    rk::GView g;
    d3 C;
    const uint32_t *pixels;
    uint32_t n;
    uint32_t chunk;  // lights per chunk (points first, then triangle lights)
    uint32_t sq;     // stratification of each triangle light (sq x sq samples)
    double *part;    // [chunk][3][n]
};

template <int STACK>
__global__ __launch_bounds__(256) void k_truth(const TruthArgs a) {
    __shared__ int32_t stack[STACK * 256];
    int32_t *stk = stack + threadIdx.x;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t chunk = blockIdx.y;
    const uint32_t p = a.pixels[i];
    const rk::GView &g = a.g;
    const d3 P = d3{g.pos[p], g.pos[g.P + p], g.pos[2 * g.P + p]};
    const rt::f3 nf = rk::load_normal(g, p);
    const float nx = nf.x, ny = nf.y, nz = nf.z;
    const d3 alb = d3{g.alb[p], g.alb[g.P + p], g.alb[2 * g.P + p]};
    const float emissive = g.alb[3 * g.P + p];
    const double rough = g.mat[p], metal = g.mat[g.P + p];
    d3 acc{0.0, 0.0, 0.0};
    if (emissive > 0.5f) {
        if (chunk == 0) acc = alb;
    } else if ((nx * nx + ny * ny) + nz * nz != 0.0f) {
        const d3 N = d3{nx, ny, nz};
        const uint32_t nP = static_cast<uint32_t>(a.sc.nPoints), nT = static_cast<uint32_t>(a.sc.nTriLights);
        const uint32_t l0 = chunk * a.chunk, l1 = min(l0 + a.chunk, nP + nT);
        for (uint32_t l = l0; l < l1; ++l) {
            if (l < nP) {
                const restir_point_light &pl = a.sc.points[l];
                const d3 L = ld3(pl.pos);
                d3 f;
                integrand(P, N, a.C, L, nullptr, alb, rough, metal, ld3(pl.color_luminance), f);
                if (f.x == 0.0 && f.y == 0.0 && f.z == 0.0) continue;
                if (occluded<STACK>(a.sc, P, L, stk)) continue;
                acc = acc + f;
            } else {
                const restir_tri_light &tl = a.sc.triLights[l - nP];
                const d3 LN = ld3(tl.normalArea), Le = ld3(tl.emission_luminance);
                const d3 q1 = ld3(tl.p1), q2 = ld3(tl.p2), q3 = ld3(tl.p3);
                d3 sum{0.0, 0.0, 0.0};
                for (uint32_t u = 0; u < a.sq; ++u)
                    for (uint32_t v = 0; v < a.sq; ++v) {
                        const double r1 = (u + 0.5) / a.sq, r2 = (v + 0.5) / a.sq, sr = sqrt(r1);
                        const d3 L = q1 * (1.0 - sr) + q2 * (sr * (1.0 - r2)) + q3 * (r2 * sr);
                        d3 f;
                        integrand(P, N, a.C, L, &LN, alb, rough, metal, Le, f);
                        if (f.x == 0.0 && f.y == 0.0 && f.z == 0.0) continue;
                        if (occluded<STACK>(a.sc, P, L, stk)) continue;
                        sum = sum + f;
                    }
                acc = acc + sum * (static_cast<double>(tl.normalArea[3]) / (static_cast<double>(a.sq) * a.sq));
            }
        }
    }
    double *out = a.part + static_cast<size_t>(chunk) * 3 * a.n;
    out[i] = acc.x;
    out[a.n + i] = acc.y;
    out[2 * static_cast<size_t>(a.n) + i] = acc.z;
}

__global__ __launch_bounds__(256) void k_truth_reduce(const double *part, uint32_t chunks, uint32_t n,
                                                      double *rgb, uint64_t stride) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (uint32_t k = 0; k < chunks; ++k) s += part[(static_cast<size_t>(k) * 3 + c) * n + i];
        rgb[c * stride + i] = s;
    }
}

}  // namespace rgt

namespace rk {

size_t truth_scratch_bytes(const SceneView &sc, uint32_t n) {
    const uint32_t lights = static_cast<uint32_t>(sc.nPoints + sc.nTriLights);
    const uint32_t chunks = lights == 0 ? 1 : (lights + kTruthChunk - 1) / kTruthChunk;
    return sizeof(double) * 3 * static_cast<size_t>(n) * chunks;
}

hipError_t launch_truth(int stack, const SceneView &sc, const GView &g, const float cam[3],
                        const uint32_t *pixels, uint32_t n, uint32_t sq, void *scratch, double *rgb,
                        uint64_t stride, hipStream_t s) {
    const uint32_t lights = static_cast<uint32_t>(sc.nPoints + sc.nTriLights);
    const uint32_t chunks = lights == 0 ? 1 : (lights + kTruthChunk - 1) / kTruthChunk;
    rgt::TruthArgs a{sc, g, rgt::d3{cam[0], cam[1], cam[2]}, pixels, n, kTruthChunk, sq,
                     static_cast<double *>(scratch)};
    const dim3 grid((n + 255) / 256, chunks);
    if (stack <= 32) RK_LAUNCH("k_truth", (rgt::k_truth<32>), grid, dim3(256), 0, s, a);
    else RK_LAUNCH("k_truth", (rgt::k_truth<64>), grid, dim3(256), 0, s, a);
    RK_LAUNCH("k_truth_reduce", rgt::k_truth_reduce, dim3((n + 255) / 256), dim3(256), 0, s,
              static_cast<const double *>(scratch), chunks, n, rgb, stride);
    return hipGetLastError();
}

}  // namespace rk
