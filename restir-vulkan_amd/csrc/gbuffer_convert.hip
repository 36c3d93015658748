// gbuffer_convert.hip — the reference's G-buffer attachments (gBufferPass.cpp:75-108) decoded into
// the planes the passes read: restir_gbuffer_from_attachments. One lane per pixel reads one texel
// of each attachment (rows at row_bytes, texels packed) and writes the 13 planes (normals as float
// or half). The decode follows the Vulkan rules for sampling each format; the sRGB transfer is a
// 256-entry table computed on the host in double precision and rounded to float.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "ctx.h"
#include "kernels.h"

using namespace rki;

namespace {

struct Att {
    const unsigned char *p;
    uint32_t fmt, row;
};

struct ConvArgs {
    Att alb, nrm, mat, pos, dep;
    uint32_t W, H;
    float *outPos;
    void *outNrm;
    bool nrmHalf;
    float *outAlb, *outMat, *outDep;
    uint64_t P;
    float srgb[256];  // sRGB8 -> linear
};

uint32_t texel_bytes(uint32_t f) {
    switch (f) {
        case RESTIR_TEXEL_RGBA8_SRGB:
        case RESTIR_TEXEL_RGBA8_UNORM:
        case RESTIR_TEXEL_RG16_UNORM:
        case RESTIR_TEXEL_D32_SFLOAT:
        case RESTIR_TEXEL_D24_UNORM: return 4;
        case RESTIR_TEXEL_RGB16_SNORM:
        case RESTIR_TEXEL_RGB16_SFLOAT: return 6;
        case RESTIR_TEXEL_RGBA16_SNORM:
        case RESTIR_TEXEL_RGBA16_SFLOAT: return 8;
        case RESTIR_TEXEL_RGB32_SFLOAT: return 12;
        case RESTIR_TEXEL_RGBA32_SFLOAT: return 16;
        default: return 0;
    }
}

__device__ __forceinline__ float snorm16(int16_t c) { return fmaxf(static_cast<float>(c) / 32767.0f, -1.0f); }
__device__ __forceinline__ float unorm16(uint16_t c) { return static_cast<float>(c) / 65535.0f; }
__device__ __forceinline__ float half_bits(uint16_t h) { return static_cast<float>(__builtin_bit_cast(_Float16, h)); }

// Four channels of texel (x, y) of `a` (missing channels read 0, as the formats here never need
// their alpha defaulted to 1).
__device__ __forceinline__ float4 fetch(const Att &a, uint32_t x, uint32_t y, const float *srgb) {
    const unsigned char *t = a.p + static_cast<size_t>(y) * a.row;
    switch (a.fmt) {
        case RESTIR_TEXEL_RGBA8_SRGB: {
            const uint32_t v = *reinterpret_cast<const uint32_t *>(t + 4u * x);
            return make_float4(srgb[v & 0xffu], srgb[(v >> 8) & 0xffu], srgb[(v >> 16) & 0xffu],
                               static_cast<float>(v >> 24) / 255.0f);
        }
        case RESTIR_TEXEL_RGBA8_UNORM: {
            const uint32_t v = *reinterpret_cast<const uint32_t *>(t + 4u * x);
            return make_float4(static_cast<float>(v & 0xffu) / 255.0f, static_cast<float>((v >> 8) & 0xffu) / 255.0f,
                               static_cast<float>((v >> 16) & 0xffu) / 255.0f, static_cast<float>(v >> 24) / 255.0f);
        }
        case RESTIR_TEXEL_RGB16_SNORM:
        case RESTIR_TEXEL_RGBA16_SNORM: {
            const int16_t *c = reinterpret_cast<const int16_t *>(t + (a.fmt == RESTIR_TEXEL_RGB16_SNORM ? 6u : 8u) * x);
            return make_float4(snorm16(c[0]), snorm16(c[1]), snorm16(c[2]), 0.0f);
        }
        case RESTIR_TEXEL_RGB16_SFLOAT:
        case RESTIR_TEXEL_RGBA16_SFLOAT: {
            const uint16_t *c = reinterpret_cast<const uint16_t *>(t + (a.fmt == RESTIR_TEXEL_RGB16_SFLOAT ? 6u : 8u) * x);
            return make_float4(half_bits(c[0]), half_bits(c[1]), half_bits(c[2]), 0.0f);
        }
        case RESTIR_TEXEL_RGB32_SFLOAT: {
            const float *c = reinterpret_cast<const float *>(t + 12u * x);
            return make_float4(c[0], c[1], c[2], 0.0f);
        }
        case RESTIR_TEXEL_RGBA32_SFLOAT: return *reinterpret_cast<const float4 *>(t + 16u * x);
        case RESTIR_TEXEL_RG16_UNORM: {
            const uint32_t v = *reinterpret_cast<const uint32_t *>(t + 4u * x);
            return make_float4(unorm16(static_cast<uint16_t>(v & 0xffffu)), unorm16(static_cast<uint16_t>(v >> 16)),
                               0.0f, 0.0f);
        }
        case RESTIR_TEXEL_D32_SFLOAT: return make_float4(*reinterpret_cast<const float *>(t + 4u * x), 0.0f, 0.0f, 0.0f);
        case RESTIR_TEXEL_D24_UNORM: {
            const uint32_t v = *reinterpret_cast<const uint32_t *>(t + 4u * x) & 0xffffffu;
            return make_float4(static_cast<float>(v) / 16777215.0f, 0.0f, 0.0f, 0.0f);
        }
        default: return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

__global__ __launch_bounds__(256) void k_gbuffer_convert(const ConvArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.W * a.H) return;
    const uint32_t x = i % a.W, y = i / a.W;
    const uint64_t P = a.P;
    const float4 al = fetch(a.alb, x, y, a.srgb), nr = fetch(a.nrm, x, y, a.srgb);
    const float4 mt = fetch(a.mat, x, y, a.srgb), ps = fetch(a.pos, x, y, a.srgb);
    const float4 dp = fetch(a.dep, x, y, a.srgb);
    a.outPos[i] = ps.x;
    a.outPos[P + i] = ps.y;
    a.outPos[2 * P + i] = ps.z;
    if (a.nrmHalf) {
        _Float16 *h = static_cast<_Float16 *>(a.outNrm);
        h[i] = static_cast<_Float16>(nr.x);
        h[P + i] = static_cast<_Float16>(nr.y);
        h[2 * P + i] = static_cast<_Float16>(nr.z);
    } else {
        float *f = static_cast<float *>(a.outNrm);
        f[i] = nr.x;
        f[P + i] = nr.y;
        f[2 * P + i] = nr.z;
    }
    a.outAlb[i] = al.x;
    a.outAlb[P + i] = al.y;
    a.outAlb[2 * P + i] = al.z;
    a.outAlb[3 * P + i] = al.w;
    a.outMat[i] = mt.x;
    a.outMat[P + i] = mt.y;
    a.outDep[i] = dp.x;
}

// the formats each attachment may have (as the reference's findSupportedFormat candidates allow)
bool allowed(int which, uint32_t f) {
    switch (which) {
        case 0: return f == RESTIR_TEXEL_RGBA8_SRGB || f == RESTIR_TEXEL_RGBA8_UNORM || f == RESTIR_TEXEL_RGBA32_SFLOAT;
        case 1: return f == RESTIR_TEXEL_RGB16_SNORM || f == RESTIR_TEXEL_RGBA16_SNORM || f == RESTIR_TEXEL_RGB16_SFLOAT ||
                       f == RESTIR_TEXEL_RGBA16_SFLOAT || f == RESTIR_TEXEL_RGB32_SFLOAT || f == RESTIR_TEXEL_RGBA32_SFLOAT;
        case 2: return f == RESTIR_TEXEL_RG16_UNORM || f == RESTIR_TEXEL_RGBA32_SFLOAT;
        case 3: return f == RESTIR_TEXEL_RGB32_SFLOAT || f == RESTIR_TEXEL_RGBA32_SFLOAT;
        default: return f == RESTIR_TEXEL_D32_SFLOAT || f == RESTIR_TEXEL_D24_UNORM;
    }
}

}  // namespace

extern "C" int restir_gbuffer_from_attachments(restir_ctx *ctx, const restir_gbuffer_attachments *att,
                                               float *position, void *normal, uint32_t normalFormat, float *albedo,
                                               float *material, float *depth, uint64_t planeStride, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (!att || !position || !normal || !albedo || !material || !depth)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "gbuffer_from_attachments: NULL argument");
    const uint32_t W = att->width, H = att->height;
    if (W == 0 || H == 0 || static_cast<uint64_t>(W) * H > (1ull << 31))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "gbuffer_from_attachments: bad size %ux%u", W, H);
    if (planeStride < static_cast<uint64_t>(W) * H)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "gbuffer_from_attachments: plane_stride < width*height");
    if (normalFormat != RESTIR_NORMAL_F32 && normalFormat != RESTIR_NORMAL_F16)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "gbuffer_from_attachments: unknown normal format %u", normalFormat);
    const restir_attachment *in[5] = {&att->albedo, &att->normal, &att->material, &att->world_position, &att->depth};
    const char *names[5] = {"albedo", "normal", "material", "world_position", "depth"};
    ConvArgs a{};
    Att *outAtt[5] = {&a.alb, &a.nrm, &a.mat, &a.pos, &a.dep};
    for (int k = 0; k < 5; ++k) {
        const restir_attachment &r = *in[k];
        if (!r.data) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "gbuffer_from_attachments: %s is NULL", names[k]);
        if (!allowed(k, r.format))
            return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "gbuffer_from_attachments: %s format %u is not accepted",
                        names[k], r.format);
        const uint32_t tb = texel_bytes(r.format);
        if (static_cast<uint64_t>(r.row_bytes) < static_cast<uint64_t>(W) * tb)
            return fail(ctx, RESTIR_E_SIZE_MISMATCH, "gbuffer_from_attachments: %s rows of %u bytes < %u texels of %u",
                        names[k], r.row_bytes, W, tb);
        // texel reads are naturally aligned words
        const uint32_t align = tb == 6 ? 2u : (tb >= 16 ? 16u : 4u);
        if (reinterpret_cast<uintptr_t>(r.data) % align != 0 || r.row_bytes % align != 0)
            return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "gbuffer_from_attachments: %s is not %u-byte aligned", names[k],
                        align);
        *outAtt[k] = Att{static_cast<const unsigned char *>(r.data), r.format, r.row_bytes};
    }
    for (int c = 0; c < 256; ++c) {
        const double v = c / 255.0;
        a.srgb[c] = static_cast<float>(v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4));
    }
    a.W = W;
    a.H = H;
    a.outPos = position;
    a.outNrm = normal;
    a.nrmHalf = normalFormat == RESTIR_NORMAL_F16;
    a.outAlb = albedo;
    a.outMat = material;
    a.outDep = depth;
    a.P = planeStride;
    if (int r = use_device(ctx)) return r;
    const uint64_t n = static_cast<uint64_t>(W) * H;
    RK_LAUNCH("k_gbuffer_convert", k_gbuffer_convert, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0,
              static_cast<hipStream_t>(stream), a);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "gbuffer_from_attachments launch");
}
