// scan.h — exclusive prefix sums over device arrays for the scene builders: 256-lane groups of
// 1024 items (4 per lane, wave64 shuffles, one LDS word per wave), then the group totals scanned
// the same way and added back.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "kernels.h"

namespace rscan {

constexpr uint32_t kItems = 1024;  // items per 256-lane group

template <typename T>
__device__ inline T wave_inclusive(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// out[i] = sum of in[j] for j < i inside each group of kItems; the group's total goes to
// sums[blockIdx.x]. in == out is allowed (each lane reads its items before writing them).
template <typename T>
__global__ __launch_bounds__(256) void k_scan_groups(const T *in, T *out, uint32_t n, T *sums) {
    __shared__ T waveTotal[4];
    const uint32_t base = blockIdx.x * kItems + threadIdx.x * 4u;
    T v[4];
    T run = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const T x = base + i < n ? in[base + i] : T(0);
        v[i] = run;
        run += x;
    }
    const T incl = wave_inclusive(run);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) waveTotal[w] = incl;
    __syncthreads();
    T off = incl - run;
    for (int k = 0; k < w; ++k) off += waveTotal[k];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (base + i < n) out[base + i] = off + v[i];
    if (threadIdx.x == 255) sums[blockIdx.x] = off + run;
}

template <typename T>
__global__ __launch_bounds__(256) void k_scan_add(T *out, uint32_t n, const T *groupOffset) {
    const T add = groupOffset[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t j = blockIdx.x * kItems + k * 256u + threadIdx.x;
        if (j < n) out[j] += add;
    }
}

// Items of temporary storage exclusive_scan needs for n items.
inline size_t temp_items(uint32_t n) {
    size_t t = 1;
    while (n > 1) {
        n = (n + kItems - 1) / kItems;
        t += n;
    }
    return t;
}

// Exclusive scan in -> out (n >= 1); the grand total is written to *total (device) when non-null.
template <typename T>
hipError_t exclusive_scan(const T *in, T *out, uint32_t n, T *temp, T *total, hipStream_t s) {
    const uint32_t groups = (n + kItems - 1) / kItems;
    T *sums = (groups == 1 && total) ? total : temp;
    RK_LAUNCH("k_scan_groups", k_scan_groups<T>, dim3(groups), dim3(256), 0, s, in, out, n, sums);
    if (groups > 1) {
        if (hipError_t e = exclusive_scan<T>(temp, temp, groups, temp + groups, total, s)) return e;
        RK_LAUNCH("k_scan_add", k_scan_add<T>, dim3(groups), dim3(256), 0, s, out, n, static_cast<const T *>(temp));
    }
    return hipGetLastError();
}

}  // namespace rscan
