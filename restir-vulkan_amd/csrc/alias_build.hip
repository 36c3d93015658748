// alias_build.hip — createAliasTable (src/misc.cpp:418-497) on the device: the reference's table,
// column for column and bit for bit, in one workgroup and one launch (no host round trip, so it
// can sit inside a captured graph).
//
// The reference sums the light powers in index order in float, normalises them to mean 1, queues
// the items >= 1 ("large") and < 1 ("small") in index order, then repeatedly pairs the front large
// g with the front small l: column l = (scaled[l], alias g), scaled[g] = (scaled[g] + scaled[l]) -
// 1, and g goes to the back of the small or the large queue. Taken one round of the large queue at
// a time (the larges queued when the round starts, in order), the pairs of a round are
// independent: the k-th large of the round meets the k-th small still queued. A round's pairs are
// therefore formed in parallel, and the larges that stay large (turn small) are appended in pair
// order through a workgroup scan. When fewer smalls are queued than the round has larges, the
// round goes in steps, each pairing as many as are queued; what a step appends to the small queue
// serves the next step. The power sum stays sequential (one lane over LDS-staged chunks): float
// addition in another order would round differently.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ctx.h"
#include "kernels.h"
#include "scan.h"

using namespace rki;

namespace {

constexpr uint32_t kLanes = 1024;  // one workgroup of 16 waves
constexpr bool kSwitchedOff = true;

// exclusive scan of v over the workgroup; *total gets the sum
__device__ inline uint32_t group_exclusive(uint32_t v, uint32_t *waveTotal, uint32_t *total) {
    const uint32_t incl = rscan::wave_inclusive(v);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) waveTotal[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t k = 0; k < kLanes / 64; ++k) {
        const uint32_t t = waveTotal[k];
        off += k < w ? t : 0u;
        tot += t;
    }
    __syncthreads();  // waveTotal is reused by the next call
    *total = tot;
    return off + incl - v;
}

__device__ inline void settle(restir_alias_table_column *cols, uint32_t i) {  // prob 1, alias itself
    cols[i].prob = 1.0f;
    cols[i].alias = static_cast<int32_t>(i);
}

// counts of a scan step packed in one word: smalls in the low 16 bits, larges in the high 16
constexpr uint32_t kSmall = 1u, kLarge = 1u << 16;

__global__ __launch_bounds__(kLanes) void k_alias_build(const restir_point_light *points, const restir_tri_light *triLights,
                                                        uint32_t n, restir_alias_table_column *cols, float *power,
                                                        float *scaled, uint32_t *under, uint32_t *overA,
                                                        uint32_t *overB) {
    __shared__ float staged[kLanes];
    __shared__ uint32_t waveTotal[kLanes / 64];
    __shared__ float shSum;
    const uint32_t tid = threadIdx.x;

    // light powers (misc.cpp:426-442) and their sum in index order
    float sum = 0.0f;
    for (uint32_t base = 0; base < n; base += kLanes) {
        const uint32_t i = base + tid;
        float w = 0.0f;
        if (i < n) {
            w = points ? points[i].color_luminance[3] : triLights[i].emission_luminance[3] * triLights[i].normalArea[3];
            power[i] = w;
        }
        staged[tid] = w;
        __syncthreads();
        if (tid == 0) {
            const uint32_t m = n - base < kLanes ? n - base : kLanes;
            for (uint32_t k = 0; k < m; ++k) sum += staged[k];
        }
        __syncthreads();
    }
    if (tid == 0) shSum = sum;
    __syncthreads();
    const float powerSum = shSum;
    const float count = static_cast<float>(n);

    // normalise; queue larges and smalls in index order (misc.cpp:446-456)
    uint32_t nLarge = 0, nSmall = 0;
    for (uint32_t base = 0; base < n; base += kLanes) {
        const uint32_t i = base + tid;
        uint32_t kind = 0;
        if (i < n) {
            const float w = power[i];
            const float s = count * w / powerSum;
            scaled[i] = s;
            restir_alias_table_column c;
            c.prob = 0.0f;
            c.alias = -1;
            c.oriProb = w / powerSum;
            c.aliasOriProb = 0.0f;
            cols[i] = c;
            kind = s >= 1.0f ? kLarge : kSmall;
        }
        uint32_t tot;
        const uint32_t ex = group_exclusive(kind, waveTotal, &tot);
        if (kind == kLarge) overA[nLarge + (ex >> 16)] = i;
        else if (kind == kSmall) under[nSmall + (ex & 0xffffu)] = i;
        nLarge += tot >> 16;
        nSmall += tot & 0xffffu;
    }

    // pairing (misc.cpp:459-477), one round of the large queue at a time
    uint32_t head = 0, tail = nSmall, m = nLarge;  // small queue is under[head, tail)
    uint32_t *cur = overA, *nxt = overB;
    bool cut = false;  // the small queue ran dry inside a round
    while (m > 0 && head < tail) {
        uint32_t k0 = 0, mNext = 0;
        while (k0 < m && head < tail) {
            const uint32_t c = m - k0 < tail - head ? m - k0 : tail - head;
            for (uint32_t b = 0; b < c; b += kLanes) {
                const uint32_t j = b + tid;
                uint32_t kind = 0, g = 0;
                if (j < c) {
                    g = cur[k0 + j];
                    const uint32_t l = under[head + j];
                    const float sl = scaled[l];
                    cols[l].prob = sl;
                    cols[l].alias = static_cast<int32_t>(g);
                    const float v = (scaled[g] + sl) - 1.0f;
                    scaled[g] = v;
                    kind = v < 1.0f ? kSmall : kLarge;
                }
                uint32_t tot;
                const uint32_t ex = group_exclusive(kind, waveTotal, &tot);
                if (kind == kSmall) under[tail + (ex & 0xffffu)] = g;
                else if (kind == kLarge) nxt[mNext + (ex >> 16)] = g;
                tail += tot & 0xffffu;
                mNext += tot >> 16;
            }
            head += c;
            k0 += c;
            __syncthreads();  // appended queue entries and new weights visible to the next step
        }
        if (k0 < m) {  // still queued as large: the rest of this round, then its survivors
            for (uint32_t j = k0 + tid; j < m; j += kLanes) settle(cols, cur[j]);
            for (uint32_t j = tid; j < mNext; j += kLanes) settle(cols, nxt[j]);
            cut = true;
            break;
        }
        uint32_t *t = cur;
        cur = nxt;
        nxt = t;
        m = mNext;
    }
    if (!cut)
        for (uint32_t j = tid; j < m; j += kLanes) settle(cols, cur[j]);  // larges left (misc.cpp:479-484)
    for (uint32_t j = head + tid; j < tail; j += kLanes) settle(cols, under[j]);  // smalls left (486-491)
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kLanes) cols[i].aliasOriProb = cols[cols[i].alias].oriProb;  // 493-495
}

}  // namespace

extern "C" int restir_build_alias_table_device(restir_ctx *ctx, const restir_point_light *points, uint32_t nPoint,
                                               const restir_tri_light *triLights, uint32_t nTri,
                                               restir_alias_table_column *columns, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    const uint32_t n = nPoint > 0 ? nPoint : nTri;
    if (n == 0) return RESTIR_OK;
    if (!columns || (nPoint > 0 && !points) || (nPoint == 0 && !triLights))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "build_alias_table_device: NULL buffer");
    if (n > (1u << 30)) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "build_alias_table_device: %u lights > 2^30", n);
    const size_t colBytes = sizeof(restir_alias_table_column) * static_cast<size_t>(n);
    const void *in = nPoint > 0 ? static_cast<const void *>(points) : static_cast<const void *>(triLights);
    const size_t inBytes = nPoint > 0 ? sizeof(restir_point_light) * static_cast<size_t>(n)
                                      : sizeof(restir_tri_light) * static_cast<size_t>(n);
    if (overlaps(columns, colBytes, in, inBytes))
        return fail(ctx, RESTIR_E_ALIASING, "build_alias_table_device: columns overlap the light records");
    // Switched off: a test run faulted the GPU while this kernel's 5000-round case ran, after
    // the same case had passed in an earlier run, and no cause has been identified (a CPU
    // emulation of every index stays in bounds and matches the host table). Until one is, the
    // entry point refuses instead of launching.
    if (kSwitchedOff)
        return fail(ctx, RESTIR_E_NOT_READY,
                    "build_alias_table_device is switched off (unexplained GPU fault under investigation); "
                    "use restir_build_alias_table on the host");
    if (int r = use_device(ctx)) return r;
    if (int r = ensure(ctx, &ctx->aliasScratch, &ctx->aliasScratchBytes, 20ull * n)) return r;
    float *power = static_cast<float *>(ctx->aliasScratch);
    float *scaled = power + n;
    uint32_t *under = reinterpret_cast<uint32_t *>(scaled + n);
    uint32_t *overA = under + n, *overB = overA + n;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    RK_LAUNCH("k_alias_build", k_alias_build, dim3(1), dim3(kLanes), 0, s, nPoint > 0 ? points : nullptr,
              nPoint > 0 ? nullptr : triLights, n, columns, power, scaled, under, overA, overB);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "build_alias_table_device launch");
}
