// abi_mlp.hip — C ABI of the learned light proposal (declared in include/restir_abi.h).
#include <hip/hip_runtime.h>

#include <vector>

#include "ctx.h"
#include "kernels.h"
#include "mlp.h"

using namespace rki;

static_assert(sizeof(restir_mlp) == 56, "restir_mlp layout (the Python binding mirrors it)");

namespace {

int check_mlp(restir_ctx *ctx, const restir_mlp *m, const char *name) {
    if (!m) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: mlp is NULL", name);
    const uint32_t K = m->n_clusters;
    if (K != 64 && K != 128 && K != 256 && (K % rm::kSuper != 0 || K > static_cast<uint32_t>(rm::kMaxClusters)))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: n_clusters %u is not 64, 128, 256 or a multiple of %d up to %d",
                    name, K, rm::kSuper, rm::kMaxClusters);
    if (!m->weights_bf16 || !m->bias)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: mlp weight buffers are NULL", name);
    if (reinterpret_cast<uintptr_t>(m->weights_bf16) % 16 != 0)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: weights_bf16 is not 16-byte aligned", name);
    return RESTIR_OK;
}

int check_box(restir_ctx *ctx, const float *lo, const float *hi, const char *name) {
    if (!lo || !hi) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: scene box is NULL", name);
    for (int k = 0; k < 3; ++k)
        if (!(hi[k] > lo[k])) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: empty scene box", name);
    return RESTIR_OK;
}

template <typename T>
int fetch(restir_ctx *ctx, std::vector<T> &dst, const T *src, size_t n, const char *what) {
    dst.resize(n);
    const hipError_t e = hipMemcpy(dst.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost);
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, what);
}

}  // namespace

extern "C" {

int restir_mlp_init(restir_ctx *ctx, restir_mlp *m, uint64_t seed, const float *prior, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (int r = check_mlp(ctx, m, "mlp_init")) return r;
    if (!m->weights || !m->adam_m || !m->adam_v || !prior)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_init: NULL buffer");
    if (int r = use_device(ctx)) return r;
    m->adam_step = 0;
    hipError_t e = rk::mlp_init(*m, seed, prior, static_cast<hipStream_t>(stream));
    if (e == hipSuccess && m->adam_step_device) {
        rk::k_zero_u32<<<1, 64, 0, static_cast<hipStream_t>(stream)>>>(m->adam_step_device, 1u);
        e = hipGetLastError();
    }
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "mlp_init launch");
}

int restir_set_light_clusters(restir_ctx *ctx, uint32_t k, const int32_t *lightCluster,
                              const float *lightPWithin, const int32_t *clusterBegin,
                              const int32_t *clusterCount, const int32_t *clusterLights,
                              const restir_alias_table_column *clusterAlias) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    ctx->clustersSet = false;
    if (!ctx->lightsSet) return fail(ctx, RESTIR_E_NOT_READY, "set_light_clusters: set lights first");
    if (!lightCluster || !lightPWithin || !clusterBegin || !clusterCount || !clusterLights || !clusterAlias)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_light_clusters: NULL buffer");
    if (k == 0 || k > static_cast<uint32_t>(ctx->scene.nLights))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_light_clusters: k=%u", k);
    if (int r = use_device(ctx)) return r;
    // validate: the kernels index lights and cluster slots with these values
    const size_t n = static_cast<size_t>(ctx->scene.nLights);  // the alias table's index space
    std::vector<int32_t> lc, beg, cnt, lights;
    std::vector<restir_alias_table_column> al;
    if (int r = fetch(ctx, lc, lightCluster, n, "set_light_clusters: light_cluster")) return r;
    if (int r = fetch(ctx, beg, clusterBegin, k, "set_light_clusters: cluster_begin")) return r;
    if (int r = fetch(ctx, cnt, clusterCount, k, "set_light_clusters: cluster_count")) return r;
    if (int r = fetch(ctx, lights, clusterLights, n, "set_light_clusters: cluster_lights")) return r;
    if (int r = fetch(ctx, al, clusterAlias, n, "set_light_clusters: cluster_alias")) return r;
    std::vector<float> pw;
    if (int r = fetch(ctx, pw, lightPWithin, n, "set_light_clusters: light_pwithin")) return r;
    std::vector<restir_alias_table_column> src;  // the global table, past its 16-byte header
    if (int r = fetch(ctx, src, reinterpret_cast<const restir_alias_table_column *>(ctx->scene.alias), n,
                      "set_light_clusters: alias table"))
        return r;
    for (size_t l = 0; l < n; ++l) {
        if (lc[l] < 0 || lc[l] >= static_cast<int32_t>(k))
            return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_light_clusters: light %zu in cluster %d", l, lc[l]);
        if (lights[l] < 0 || lights[l] >= static_cast<int32_t>(n))
            return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_light_clusters: slot %zu names light %d", l, lights[l]);
    }
    for (uint32_t c = 0; c < k; ++c) {
        if (beg[c] < 0 || cnt[c] < 0 || static_cast<size_t>(beg[c]) + cnt[c] > n)
            return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_light_clusters: cluster %u range", c);
        for (int32_t i = 0; i < cnt[c]; ++i)
            if (al[beg[c] + i].alias < 0 || al[beg[c] + i].alias >= cnt[c])
                return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_light_clusters: cluster %u alias", c);
        // The proposal's density q(l) reads p_within from whichever record is at hand (the column
        // that drew l, or light_pwithin[l]) and p_src likewise: they must be the same numbers
        for (int32_t i = 0; i < cnt[c]; ++i) {
            const restir_alias_table_column &col = al[beg[c] + i];
            const int32_t l = lights[beg[c] + i];
            if (lc[l] != static_cast<int32_t>(c))
                return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_light_clusters: light %d listed in cluster %u but assigned %d", l, c, lc[l]);
            if (pw[l] != col.oriProb || col.aliasOriProb != al[beg[c] + col.alias].oriProb)
                return fail(ctx, RESTIR_E_INVALID_ARGUMENT,
                            "set_light_clusters: cluster %u: light_pwithin / oriProb / aliasOriProb disagree", c);
        }
    }
    for (size_t l = 0; l < n; ++l)
        if (src[l].aliasOriProb != src[src[l].alias].oriProb)
            return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_light_clusters: alias table column %zu: aliasOriProb is not oriProb of its alias", l);
    ctx->nClusters = static_cast<int32_t>(k);
    ctx->lightCluster = lightCluster;
    ctx->lightPWithin = lightPWithin;
    ctx->clusterBegin = clusterBegin;
    ctx->clusterCount = clusterCount;
    ctx->clusterLights = clusterLights;
    ctx->clusterAlias = clusterAlias;
    // the fused slot records k_propose reads: one 32-byte load per drawn slot gives the light, its
    // p_within and its p_src
    std::vector<float4> fused(2 * n);
    for (uint32_t c = 0; c < k; ++c)
        for (int32_t i = 0; i < cnt[c]; ++i) {
            const size_t slot = static_cast<size_t>(beg[c]) + i;
            const restir_alias_table_column &col = al[slot];
            const int32_t lc = lights[slot], la = lights[beg[c] + col.alias];
            fused[2 * slot] = make_float4(col.prob, __builtin_bit_cast(float, lc), __builtin_bit_cast(float, la), col.oriProb);
            fused[2 * slot + 1] = make_float4(col.aliasOriProb, src[lc].oriProb, src[la].oriProb, 0.0f);
        }
    if (int r = ensure(ctx, &ctx->clusterFused, &ctx->clusterFusedBytes, sizeof(float4) * fused.size())) return r;
    const hipError_t e = hipMemcpy(ctx->clusterFused, fused.data(), sizeof(float4) * fused.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(ctx, e, "set_light_clusters: fused slot records");
    ctx->clustersSet = true;
    return RESTIR_OK;
}

int restir_learned_candidates(restir_ctx *ctx, const restir_uniforms *u, const restir_gbuffer *g,
                              const restir_mlp *m, const float lo[3], const float hi[3], uint32_t M,
                              float mix, int32_t *candLight, float *candQ, void *stream) {
    if (int r = check_frame(ctx, u, "learned_candidates")) return r;
    if (int r = check_mlp(ctx, m, "learned_candidates")) return r;
    if (int r = check_box(ctx, lo, hi, "learned_candidates")) return r;
    if (!ctx->clustersSet) return fail(ctx, RESTIR_E_NOT_READY, "learned_candidates: light clusters not set");
    if (static_cast<int32_t>(m->n_clusters) != ctx->nClusters)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "learned_candidates: mlp has %u clusters, tables %d",
                    m->n_clusters, ctx->nClusters);
    if (M < 1 || M > 16) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "learned_candidates: M=%u outside [1,16]", M);
    if (m->n_clusters > 256 && M > static_cast<uint32_t>(rm::kWideCand))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "learned_candidates: M=%u above %d with %u clusters", M,
                    rm::kWideCand, m->n_clusters);
    if (!(mix > 0.0f && mix <= 1.0f))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "learned_candidates: mix must be in (0, 1]");
    if (!candLight || !candQ) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "learned_candidates: NULL output");
    const uint32_t W = u->screenSize[0], H = u->screenSize[1];
    if (int r = check_gbuffer(ctx, g, W, H, false, "learned_candidates")) return r;
    if (int r = use_device(ctx)) return r;
    const rt::f3 cam{u->cameraPos[0], u->cameraPos[1], u->cameraPos[2]};
    const uint32_t seed = u->frame * 0x9e3779b1u + 0x7f4a7c15u;
    const hipError_t e = rk::mlp_propose(*m, view_of(g), cam, lo, hi, ctx->scene, ctx->lightCluster,
                                         ctx->lightPWithin, ctx->clusterBegin, ctx->clusterCount,
                                         ctx->clusterLights, ctx->clusterAlias,
                                         static_cast<const float4 *>(ctx->clusterFused), M, mix, seed, candLight,
                                         candQ, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "learned_candidates launch");
}

int restir_initial_temporal_candidates(restir_ctx *ctx, const restir_uniforms *u, const restir_gbuffer *cur,
                                       const restir_gbuffer *prev, const restir_reservoir *prevRes,
                                       restir_reservoir *out, const int32_t *candLight, const float *candQ,
                                       const restir_launch *launch, void *stream) {
    if (int r = check_frame(ctx, u, "initial_temporal_candidates")) return r;
    if (!ctx->lightsSet) return fail(ctx, RESTIR_E_NOT_READY, "initial_temporal_candidates: lights not set");
    const bool vis = (u->flags & RESTIR_VISIBILITY_REUSE_FLAG) != 0;
    const bool temporal = (u->flags & RESTIR_TEMPORAL_REUSE_FLAG) != 0;
    if (vis && !ctx->bvhSet) return fail(ctx, RESTIR_E_NOT_READY, "initial_temporal_candidates: BVH not set");
    if (!candLight || !candQ) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "initial_temporal_candidates: NULL candidates");
    if (u->initialLightSampleCount > 16)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "initial_temporal_candidates: at most 16 candidates");
    const uint32_t W = u->screenSize[0], H = u->screenSize[1];
    if (int r = check_gbuffer(ctx, cur, W, H, false, "initial_temporal_candidates cur")) return r;
    if (!out) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "initial_temporal_candidates: output is NULL");
    const size_t bytes = sizeof(restir_reservoir) * W * H;
    if (temporal) {
        if (int r = check_gbuffer(ctx, prev, W, H, false, "initial_temporal_candidates prev")) return r;
        if (!prevRes) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "initial_temporal_candidates: previous reservoirs are NULL");
        if (overlaps(prevRes, bytes, out, bytes))
            return fail(ctx, RESTIR_E_ALIASING, "initial_temporal_candidates: output aliases the previous reservoirs");
    }
    int tile;
    if (int r = pick_tile(ctx, launch, &tile)) return r;
    void *queues = nullptr;
    if (int r = pass_a_queues(ctx, launch, W * H, &queues)) return r;
    const hipError_t e = rk::launch_pass_a(tile, ctx->stack, *u, ctx->scene, view_of(cur), view_of(prev), prevRes,
                                           out, candLight, candQ, queues, ctx->persistentBlocks,
                                           static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "initial_temporal_candidates launch");
}

int restir_mlp_train(restir_ctx *ctx, restir_mlp *m, const restir_uniforms *u, const restir_gbuffer *g,
                     const restir_reservoir *res, const float lo[3], const float hi[3], uint32_t n,
                     float lr, uint32_t seed, float *lossOut, void *stream) {
    if (int r = check_frame(ctx, u, "mlp_train")) return r;
    if (int r = check_mlp(ctx, m, "mlp_train")) return r;
    if (int r = check_box(ctx, lo, hi, "mlp_train")) return r;
    if (!m->weights || !m->adam_m || !m->adam_v) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_train: NULL buffer");
    if (!ctx->clustersSet) return fail(ctx, RESTIR_E_NOT_READY, "mlp_train: light clusters not set");
    if (static_cast<int32_t>(m->n_clusters) != ctx->nClusters)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "mlp_train: cluster count mismatch");
    if (n == 0 || n % 64 != 0 || n > 65536)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_train: n_samples=%u (multiple of 64, <= 65536)", n);
    if (!res) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_train: reservoirs are NULL");
    const uint32_t W = u->screenSize[0], H = u->screenSize[1];
    if (int r = check_gbuffer(ctx, g, W, H, false, "mlp_train")) return r;
    if (int r = use_device(ctx)) return r;
    if (int r = ensure(ctx, &ctx->mlpScratch, &ctx->mlpScratchBytes, rk::mlp_scratch_bytes(n, m->n_clusters)))
        return r;
    const rt::f3 cam{u->cameraPos[0], u->cameraPos[1], u->cameraPos[2]};
    const hipError_t e = rk::mlp_train(*m, view_of(g), cam, lo, hi, res, ctx->lightCluster, ctx->scene.nLights,
                                       ctx->scene.mixed ? ctx->scene.nPoints : (ctx->scene.nPoints == 0 ? 0 : -1),
                                       n, lr, seed, lossOut,
                                       ctx->mlpScratch, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "mlp_train launch");
}

int restir_mlp_features(restir_ctx *ctx, const restir_gbuffer *g, const float cam[3], const float lo[3],
                        const float hi[3], const uint32_t *pixels, uint32_t n, uint16_t *out, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (!g || !cam || !pixels || !out) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_features: NULL argument");
    if (int r = check_box(ctx, lo, hi, "mlp_features")) return r;
    if (int r = check_gbuffer(ctx, g, g->width, g->height, false, "mlp_features")) return r;
    if (n == 0) return RESTIR_OK;
    if (int r = use_device(ctx)) return r;
    const hipError_t e = rk::mlp_features(view_of(g), rt::f3{cam[0], cam[1], cam[2]}, lo, hi, pixels, n, out,
                                          static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "mlp_features launch");
}

int restir_mlp_forward(restir_ctx *ctx, const restir_mlp *m, const uint16_t *X, uint32_t n, float *logits,
                       void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (int r = check_mlp(ctx, m, "mlp_forward")) return r;
    if (!X || !logits) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_forward: NULL argument");
    if (reinterpret_cast<uintptr_t>(X) % 16 != 0)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_forward: features are not 16-byte aligned");
    if (n == 0 || n % 64 != 0 || n > 65536)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_forward: n=%u (multiple of 64, <= 65536)", n);
    if (int r = use_device(ctx)) return r;
    if (int r = ensure(ctx, &ctx->mlpScratch, &ctx->mlpScratchBytes, rk::mlp_scratch_bytes(n, m->n_clusters)))
        return r;
    const hipError_t e = rk::mlp_forward(*m, X, n, logits, ctx->mlpScratch, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "mlp_forward launch");
}

int restir_mlp_loss_grad(restir_ctx *ctx, const restir_mlp *m, const uint16_t *X, const int32_t *target,
                         const uint32_t *mask, uint32_t n, float *gw, float *gb, float *loss, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (int r = check_mlp(ctx, m, "mlp_loss_grad")) return r;
    if (!X || !target || !mask || !gw || !gb || !loss)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_loss_grad: NULL argument");
    if (reinterpret_cast<uintptr_t>(X) % 16 != 0)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_loss_grad: features are not 16-byte aligned");
    if (n == 0 || n % 64 != 0 || n > 65536)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_loss_grad: n=%u (multiple of 64, <= 65536)", n);
    if (int r = use_device(ctx)) return r;
    if (int r = ensure(ctx, &ctx->mlpScratch, &ctx->mlpScratchBytes, rk::mlp_scratch_bytes(n, m->n_clusters)))
        return r;
    const hipError_t e = rk::mlp_loss_grad(*m, X, target, mask, n, gw, gb, loss, ctx->mlpScratch,
                                           static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "mlp_loss_grad launch");
}

int restir_mlp_adam(restir_ctx *ctx, restir_mlp *m, const float *gw, const float *gb, float lr, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (int r = check_mlp(ctx, m, "mlp_adam")) return r;
    if (!m->weights || !m->adam_m || !m->adam_v || !gw || !gb)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_adam: NULL buffer");
    if (!(lr >= 0.0f)) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "mlp_adam: learning rate must be >= 0");
    if (int r = use_device(ctx)) return r;
    const hipError_t e = rk::mlp_adam(*m, gw, gb, nullptr, lr, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "mlp_adam launch");
}

}  // extern "C"
