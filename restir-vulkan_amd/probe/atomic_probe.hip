// atomic_probe.hip — cost of device-wide atomics by address spread: one atomicAdd per wave
// (lane 0 of each 64-lane wave) on one word, on two words of one 128-byte line, and on 256
// different lines (16,384 words), against the same grid with no atomic; grids of 32 K, 131 K and
// 524 K waves, best of 5 launches each, timed with HIP events. Prints ms per launch and ns per
// wave atomic. Used to size the ray queues' counters: atomics on one address are served one at
// a time, atomics on different lines in parallel.
#include <hip/hip_runtime.h>
#include <cstdio>

__global__ void k_same(unsigned *c) {
    if ((threadIdx.x & 63u) == 0u) atomicAdd(&c[0], 1u);
}
__global__ void k_two(unsigned *c) {
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(&c[0], 1u);
        atomicAdd(&c[1], 1u);
    }
}
__global__ void k_spread(unsigned *c) {
    if ((threadIdx.x & 63u) == 0u) atomicAdd(&c[(blockIdx.x & 255u) * 64u], 1u);
}
__global__ void k_none(unsigned *c) {
    if ((threadIdx.x & 63u) == 0u && blockIdx.x == 0xffffffffu) c[0] = 1u;
}

#define CK(x)                                                                                     \
    do {                                                                                          \
        const hipError_t e = (x);                                                                 \
        if (e != hipSuccess) {                                                                    \
            std::printf("error %s at line %d\n", hipGetErrorString(e), __LINE__);                 \
            return 1;                                                                             \
        }                                                                                         \
    } while (0)

int main() {
    constexpr size_t kWords = 256 * 64 + 1024;  // k_spread touches words 0 .. 255 * 64
    unsigned *c = nullptr;
    CK(hipMalloc(&c, kWords * sizeof(unsigned)));
    CK(hipMemset(c, 0, kWords * sizeof(unsigned)));
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    const char *names[] = {"none", "one word", "two words of a line", "256 lines"};
    for (const int blocks : {8192, 32768, 131072}) {
        for (int kind = 0; kind < 4; ++kind) {
            float best = 1e30f;
            for (int rep = 0; rep < 5; ++rep) {
                CK(hipEventRecord(a));
                if (kind == 0) k_none<<<blocks, 256>>>(c);
                if (kind == 1) k_same<<<blocks, 256>>>(c);
                if (kind == 2) k_two<<<blocks, 256>>>(c);
                if (kind == 3) k_spread<<<blocks, 256>>>(c);
                CK(hipGetLastError());
                CK(hipEventRecord(b));
                CK(hipEventSynchronize(b));
                float ms = 0.0f;
                CK(hipEventElapsedTime(&ms, a, b));
                if (ms < best) best = ms;
            }
            const double atomics = 4.0 * blocks * (kind == 2 ? 2 : 1);
            std::printf("waves %7d  %-20s %8.3f ms  %6.2f ns per wave atomic\n", blocks * 4, names[kind], best,
                        kind ? best * 1e6 / atomics : 0.0);
        }
    }
    CK(hipFree(c));
    return 0;
}
