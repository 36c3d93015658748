// alias_kernel.h — the body of the device alias-table build (createAliasTable, src/misc.cpp:418-497),
// written once over a workgroup policy G so that the same code runs as k_alias_build (one
// 1024-lane workgroup, alias_build.hip) and in a CPU build with one thread per lane
// (tests/cpu_kernels/alias_build_cpu.cpp), where ThreadSanitizer and AddressSanitizer watch
// every access between the barriers.
//
// G provides: tid() (lane in the workgroup), sync() (workgroup barrier that also orders memory)
// and shfl_up(v, d) (the value of lane tid - d of the caller's 64-lane wave; every lane of the
// wave calls it together). The includer defines AK_FN (the function qualifiers) before
// including this file.
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
// serves the next step. The power sum stays sequential (one lane over staged chunks): float
// addition in another order would round differently.
#pragma once

#include <cstdint>

#include "../../include/restir_abi.h"

namespace ak {

constexpr uint32_t kLanes = 1024;  // one workgroup of 16 waves
// counts of a scan step packed in one word: smalls in the low 16 bits, larges in the high 16
constexpr uint32_t kSmall = 1u, kLarge = 1u << 16;

struct Shared {
    float staged[kLanes];
    uint32_t waveTotal[kLanes / 64];
    float sum;
};

template <class G>
AK_FN uint32_t wave_inclusive(G &g, uint32_t v) {
    const uint32_t lane = g.tid() & 63u;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = g.shfl_up(v, d);
        if (lane >= d) v += o;
    }
    return v;
}

// exclusive scan of v over the workgroup; *total gets the sum
template <class G>
AK_FN uint32_t group_exclusive(G &g, Shared &sh, uint32_t v, uint32_t *total) {
    const uint32_t incl = wave_inclusive(g, v);
    const uint32_t w = g.tid() >> 6;
    if ((g.tid() & 63u) == 63u) sh.waveTotal[w] = incl;
    g.sync();
    uint32_t off = 0, tot = 0;
    for (uint32_t k = 0; k < kLanes / 64; ++k) {
        const uint32_t t = sh.waveTotal[k];
        off += k < w ? t : 0u;
        tot += t;
    }
    g.sync();  // waveTotal is reused by the next call
    *total = tot;
    return off + incl - v;
}

AK_FN void settle(restir_alias_table_column *cols, uint32_t i) {  // prob 1, alias itself
    cols[i].prob = 1.0f;
    cols[i].alias = static_cast<int32_t>(i);
}

template <class G>
AK_FN void build(G &g, Shared &sh, const restir_point_light *points, const restir_tri_light *triLights, uint32_t n,
                 restir_alias_table_column *cols, float *power, float *scaled, uint32_t *under, uint32_t *overA,
                 uint32_t *overB) {
    const uint32_t tid = g.tid();

    // light powers (misc.cpp:426-442) and their sum in index order
    float sum = 0.0f;
    for (uint32_t base = 0; base < n; base += kLanes) {
        const uint32_t i = base + tid;
        float w = 0.0f;
        if (i < n) {
            w = points ? points[i].color_luminance[3] : triLights[i].emission_luminance[3] * triLights[i].normalArea[3];
            power[i] = w;
        }
        sh.staged[tid] = w;
        g.sync();
        if (tid == 0) {
            const uint32_t m = n - base < kLanes ? n - base : kLanes;
            for (uint32_t k = 0; k < m; ++k) sum += sh.staged[k];
        }
        g.sync();
    }
    if (tid == 0) sh.sum = sum;
    g.sync();
    const float powerSum = sh.sum;
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
        const uint32_t ex = group_exclusive(g, sh, kind, &tot);
        if (kind == kLarge) overA[nLarge + (ex >> 16)] = i;
        else if (kind == kSmall) under[nSmall + (ex & 0xffffu)] = i;
        nLarge += tot >> 16;
        nSmall += tot & 0xffffu;
    }

    // The queues, scaled[] and cols[] written above must be visible before any lane pairs: the
    // lane that pairs large k reads the queue entry another lane wrote after the scan's last
    // barrier (without this barrier, lane 0 could read overA[0] before its writer had stored it).
    g.sync();

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
                uint32_t kind = 0, gl = 0;
                if (j < c) {
                    gl = cur[k0 + j];
                    const uint32_t l = under[head + j];
                    const float sl = scaled[l];
                    cols[l].prob = sl;
                    cols[l].alias = static_cast<int32_t>(gl);
                    const float v = (scaled[gl] + sl) - 1.0f;
                    scaled[gl] = v;
                    kind = v < 1.0f ? kSmall : kLarge;
                }
                uint32_t tot;
                const uint32_t ex = group_exclusive(g, sh, kind, &tot);
                if (kind == kSmall) under[tail + (ex & 0xffffu)] = gl;
                else if (kind == kLarge) nxt[mNext + (ex >> 16)] = gl;
                tail += tot & 0xffffu;
                mNext += tot >> 16;
            }
            head += c;
            k0 += c;
            g.sync();  // appended queue entries and new weights visible to the next step
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
    g.sync();
    for (uint32_t i = tid; i < n; i += kLanes) cols[i].aliasOriProb = cols[cols[i].alias].oriProb;  // 493-495
}

}  // namespace ak
