// profile.hip — per-kernel timing with HIP events (restir_profile_enable / restir_profile_read)
// and the ray export of the queued Pass C (restir_unbiased_rays).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"

namespace rk {

thread_local Recorder *tl_recorder = nullptr;

hipEvent_t Recorder::take() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

void Recorder::clear() {
    for (const Span &s : spans) {
        pool.push_back(s.start);
        pool.push_back(s.stop);
    }
    spans.clear();
}

Recorder::~Recorder() {
    clear();
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

void rec_pair(const char *name, hipStream_t stream, hipEvent_t *start, hipEvent_t *stop) {
    Recorder *r = tl_recorder;
    // launches recorded into a graph are not timed (their events would belong to the graph)
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &st) != hipSuccess || st != hipStreamCaptureStatusNone) return;
    hipEvent_t a = r->take();
    hipEvent_t b = a ? r->take() : nullptr;
    if (!a || !b) {  // out of events: launch untimed rather than fail the launch
        if (a) r->pool.push_back(a);
        return;
    }
    r->spans.push_back(Recorder::Span{name, a, b});
    *start = a;
    *stop = b;
}

}  // namespace rk

using namespace rki;

extern "C" {

int restir_profile_enable(restir_ctx *ctx, int enable) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    ctx->profiling = enable != 0;
    if (rk::tl_recorder == &ctx->recorder && !ctx->profiling) rk::tl_recorder = nullptr;
    return RESTIR_OK;
}

int restir_profile_read(restir_ctx *ctx, restir_kernel_time *out, uint32_t capacity,
                        uint32_t *count) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (!count || (capacity > 0 && !out))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "profile_read: NULL output");
    if (int r = use_device(ctx)) return r;
    std::vector<restir_kernel_time> rows;
    hipError_t err = hipSuccess;
    for (const rk::Recorder::Span &s : ctx->recorder.spans) {
        float ms = 0.0f;
        hipError_t e = hipEventSynchronize(s.stop);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, s.start, s.stop);
        if (e != hipSuccess) {
            err = e;
            break;
        }
        restir_kernel_time *row = nullptr;
        for (restir_kernel_time &t : rows)
            if (std::strncmp(t.name, s.name, sizeof t.name - 1) == 0) row = &t;
        if (!row) {
            restir_kernel_time t{};
            std::strncpy(t.name, s.name, sizeof t.name - 1);
            rows.push_back(t);
            row = &rows.back();
        }
        row->launches += 1;
        row->total_ms += ms;
        if (ms > row->max_ms) row->max_ms = ms;
    }
    ctx->recorder.clear();
    if (err != hipSuccess) return hip_fail(ctx, err, "profile_read");
    *count = static_cast<uint32_t>(rows.size());
    for (uint32_t i = 0; i < rows.size() && i < capacity; ++i) out[i] = rows[i];
    return RESTIR_OK;
}

int restir_unbiased_rays(restir_ctx *ctx, int which, float *p1, float *p2, uint32_t capacity,
                         uint32_t *count, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (which != 0 && which != 1)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "unbiased_rays: which=%d is not 0 or 1", which);
    if (!count || (capacity > 0 && (!p1 || !p2)))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "unbiased_rays: NULL output");
    if (ctx->queuedNpx == 0)
        return fail(ctx, RESTIR_E_NOT_READY, "unbiased_rays: no queued Pass C has run on this context");
    if (int r = use_device(ctx)) return r;
    const hipError_t e = rk::pass_c_rays(ctx->scratch, ctx->queuedNpx, ctx->queuedK, ctx->queuedOut, which,
                                         reinterpret_cast<float4 *>(p1), reinterpret_cast<float4 *>(p2),
                                         capacity, count, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "unbiased_rays");
}

int restir_set_ray_export(restir_ctx *ctx, int on) {
    (void)on;
    return ctx ? RESTIR_OK : RESTIR_E_INVALID_ARGUMENT;
}

}  // extern "C"
