// abi_views.hip — batched-view entry points (restir_*_views): per-view validation with the same
// rules as the single-view entry points, then one persistent launch per pass.
#include <hip/hip_runtime.h>

#include <vector>

#include "ctx.h"
#include "kernels.h"

using namespace rki;

namespace {

struct Span {
    const void *p;
    size_t bytes;
};

// every output span against every other span (outputs and inputs of all views)
int check_disjoint(restir_ctx *ctx, const std::vector<Span> &outs, const std::vector<Span> &ins,
                   const char *name) {
    for (size_t i = 0; i < outs.size(); ++i) {
        for (size_t j = 0; j < outs.size(); ++j)
            if (i != j && overlaps(outs[i].p, outs[i].bytes, outs[j].p, outs[j].bytes))
                return fail(ctx, RESTIR_E_ALIASING, "%s: outputs of two views overlap", name);
        for (const Span &in : ins)
            if (overlaps(outs[i].p, outs[i].bytes, in.p, in.bytes))
                return fail(ctx, RESTIR_E_ALIASING, "%s: an output overlaps an input", name);
    }
    return RESTIR_OK;
}

int check_views(restir_ctx *ctx, const restir_view *views, uint32_t n, const char *name) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (n == 0) return RESTIR_OK;
    if (!views) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: views is NULL", name);
    if (n > RESTIR_MAX_VIEWS) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: %u views > %d", name, n, RESTIR_MAX_VIEWS);
    for (uint32_t v = 0; v < n; ++v) {
        const restir_ctx *sc = views[v].scene;
        if (!sc) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: view %u has no scene context", name, v);
        if (sc->device != ctx->device)
            return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: view %u's scene is on device %d, not %d", name, v,
                        sc->device, ctx->device);
        if (!sc->lightsSet) return fail(ctx, RESTIR_E_NOT_READY, "%s: view %u: lights not set", name, v);
    }
    return RESTIR_OK;
}

int prepare(restir_ctx *ctx, uint32_t n) {
    if (int r = use_device(ctx)) return r;
    return ensure(ctx, &ctx->viewTable, &ctx->viewTableBytes, rk::views_table_bytes(n));
}

uint32_t *view_counter(restir_ctx *ctx) { return ctx->counters + 8; }

}  // namespace

extern "C" {

int restir_initial_temporal_views(restir_ctx *ctx, const restir_view *views, uint32_t n,
                                  const restir_launch *launch, void *stream) {
    if (int r = check_views(ctx, views, n, "initial_temporal_views")) return r;
    if (n == 0) return RESTIR_OK;
    int tile;
    if (int r = pick_tile(ctx, launch, &tile)) return r;
    std::vector<rk::ViewA> rec(n);
    std::vector<Span> outs, ins;
    int stack = 24;  // the smallest compiled stack; raised to the deepest scene's
    for (uint32_t v = 0; v < n; ++v) {
        const restir_view &w = views[v];
        const restir_uniforms *u = w.uniforms;
        if (int r = check_frame(ctx, u, "initial_temporal_views")) return r;
        const bool vis = (u->flags & RESTIR_VISIBILITY_REUSE_FLAG) != 0;
        const bool temporal = (u->flags & RESTIR_TEMPORAL_REUSE_FLAG) != 0;
        if (vis && !w.scene->bvhSet) return fail(ctx, RESTIR_E_NOT_READY, "initial_temporal_views: view %u: BVH not set", v);
        const uint32_t W = u->screenSize[0], H = u->screenSize[1];
        if (int r = check_gbuffer(ctx, w.cur, W, H, false, "initial_temporal_views cur")) return r;
        if (!w.out) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "initial_temporal_views: view %u output is NULL", v);
        const size_t bytes = sizeof(restir_reservoir) * W * H;
        if (temporal) {
            if (int r = check_gbuffer(ctx, w.prev, W, H, false, "initial_temporal_views prev")) return r;
            if (!w.in) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "initial_temporal_views: view %u previous reservoirs are NULL", v);
            ins.push_back({w.in, bytes});
        }
        outs.push_back({w.out, bytes});
        if (w.scene->stack > stack) stack = w.scene->stack;
        rec[v] = rk::ViewA{*u, w.scene->scene, view_of(w.cur), view_of(temporal ? w.prev : nullptr), w.in, w.out};
    }
    if (int r = check_disjoint(ctx, outs, ins, "initial_temporal_views")) return r;
    if (int r = prepare(ctx, n)) return r;
    const hipError_t e = rk::launch_pass_a_views(tile, stack, rec.data(), n, ctx->viewTable, view_counter(ctx),
                                                 resident_blocks(ctx->cus, static_cast<uint32_t>(stack) * 1024u), static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "initial_temporal_views launch");
}

int restir_spatial_views(restir_ctx *ctx, const restir_view *views, uint32_t n, int iter,
                         const restir_launch *launch, void *stream) {
    if (int r = check_views(ctx, views, n, "spatial_views")) return r;
    if (n == 0) return RESTIR_OK;
    int tile;
    if (int r = pick_tile(ctx, launch, &tile)) return r;
    std::vector<rk::ViewB> rec(n);
    std::vector<Span> outs, ins;
    for (uint32_t v = 0; v < n; ++v) {
        const restir_view &w = views[v];
        const restir_uniforms *u = w.uniforms;
        if (int r = check_frame(ctx, u, "spatial_views")) return r;
        const uint32_t W = u->screenSize[0], H = u->screenSize[1];
        if (int r = check_gbuffer(ctx, w.cur, W, H, true, "spatial_views")) return r;
        if (!w.in || !w.out) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "spatial_views: view %u reservoir buffer is NULL", v);
        const size_t bytes = sizeof(restir_reservoir) * W * H;
        ins.push_back({w.in, bytes});
        outs.push_back({w.out, bytes});
        rec[v] = rk::ViewB{*u, view_of(w.cur), w.in, w.out, rt::cos_degrees(u->spatialNormalThreshold)};
    }
    if (int r = check_disjoint(ctx, outs, ins, "spatial_views")) return r;
    if (int r = prepare(ctx, n)) return r;
    const hipError_t e = rk::launch_pass_b_views(tile, rec.data(), n, iter, ctx->viewTable, view_counter(ctx),
                                                 ctx->persistentBlocks, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "spatial_views launch");
}

int restir_unbiased_views(restir_ctx *ctx, const restir_view *views, uint32_t n, int k, int iter,
                          const restir_launch *launch, void *stream) {
    if (int r = check_views(ctx, views, n, "unbiased_views")) return r;
    if (n == 0) return RESTIR_OK;
    if (k < 1 || k > 16) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "unbiased_views: k=%d outside [1,16]", k);
    int tile;
    if (int r = pick_tile(ctx, launch, &tile)) return r;
    std::vector<rk::ViewC> rec(n);
    std::vector<Span> outs, ins;
    int stack = 24;  // the smallest compiled stack; raised to the deepest scene's
    for (uint32_t v = 0; v < n; ++v) {
        const restir_view &w = views[v];
        const restir_uniforms *u = w.uniforms;
        if (int r = check_frame(ctx, u, "unbiased_views")) return r;
        const bool vis = (u->flags & RESTIR_VISIBILITY_REUSE_FLAG) != 0;
        if (vis && !w.scene->bvhSet) return fail(ctx, RESTIR_E_NOT_READY, "unbiased_views: view %u: BVH not set", v);
        const uint32_t W = u->screenSize[0], H = u->screenSize[1];
        if (int r = check_gbuffer(ctx, w.cur, W, H, false, "unbiased_views")) return r;
        if (!w.in || !w.out) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "unbiased_views: view %u reservoir buffer is NULL", v);
        const size_t bytes = sizeof(restir_reservoir) * W * H;
        ins.push_back({w.in, bytes});
        outs.push_back({w.out, bytes});
        if (w.scene->stack > stack) stack = w.scene->stack;
        rec[v] = rk::ViewC{*u, w.scene->scene, view_of(w.cur), w.in, w.out};
    }
    if (int r = check_disjoint(ctx, outs, ins, "unbiased_views")) return r;
    if (int r = prepare(ctx, n)) return r;
    const hipError_t e = rk::launch_pass_c_views(tile, stack, rec.data(), n, k, iter, ctx->viewTable,
                                                 view_counter(ctx), resident_blocks(ctx->cus, static_cast<uint32_t>(stack) * 1024u),
                                                 static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "unbiased_views launch");
}

int restir_shade_views(restir_ctx *ctx, const restir_view *views, uint32_t n, float gamma,
                       void *stream) {
    if (int r = check_views(ctx, views, n, "shade_views")) return r;
    if (n == 0) return RESTIR_OK;
    if (!(gamma > 0.0f)) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "shade_views: gamma must be > 0");
    std::vector<rk::ViewShade> rec(n);
    std::vector<Span> outs, ins;
    for (uint32_t v = 0; v < n; ++v) {
        const restir_view &w = views[v];
        const restir_uniforms *u = w.uniforms;
        if (int r = check_frame(ctx, u, "shade_views")) return r;
        const uint32_t W = u->screenSize[0], H = u->screenSize[1];
        if (int r = check_gbuffer(ctx, w.cur, W, H, false, "shade_views")) return r;
        if (!w.in || !w.rgb) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "shade_views: view %u needs in and rgb", v);
        if (w.rgb_stride < static_cast<uint64_t>(W) * H)
            return fail(ctx, RESTIR_E_SIZE_MISMATCH, "shade_views: view %u rgb plane stride too small", v);
        ins.push_back({w.in, sizeof(restir_reservoir) * W * H});
        outs.push_back({w.rgb, sizeof(float) * (2 * w.rgb_stride + static_cast<uint64_t>(W) * H)});
        rec[v] = rk::ViewShade{w.scene->scene, rt::f3{u->cameraPos[0], u->cameraPos[1], u->cameraPos[2]},
                               view_of(w.cur), w.in, w.rgb, w.rgb_stride};
    }
    if (int r = check_disjoint(ctx, outs, ins, "shade_views")) return r;
    if (int r = prepare(ctx, n)) return r;
    const hipError_t e = rk::launch_shade_views(rec.data(), n, gamma, ctx->viewTable, view_counter(ctx),
                                                ctx->persistentBlocks, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "shade_views launch");
}

}  // extern "C"
