// valu_probe.hip — fp32 VALU issue rate of the device for three instruction mixes: eight
// independent chains per lane of (mul, add), of fma, and of (min, add), all non-packed (built with
// packed fp32 disabled like the library), 4,096 workgroups of 256 lanes, one store per lane.
// Prints lane-ops/s = VALU instructions x 64 lanes per second. Used to check the fp32 lane-op
// peak that the traversal roofline assumes (16 lanes per SIMD per clock for FMA-class ops).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
template <int OP>
__global__ __launch_bounds__(256) void k_peak(float *out, int iters, float a, float b) {
    float x0 = threadIdx.x, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (OP == 0) { x0 = x0 * a + b; x1 = x1 * a + b; x2 = x2 * a + b; x3 = x3 * a + b; x4 = x4 * a + b; x5 = x5 * a + b; x6 = x6 * a + b; x7 = x7 * a + b; }
            if (OP == 1) { x0 = fmaf(x0, a, b); x1 = fmaf(x1, a, b); x2 = fmaf(x2, a, b); x3 = fmaf(x3, a, b); x4 = fmaf(x4, a, b); x5 = fmaf(x5, a, b); x6 = fmaf(x6, a, b); x7 = fmaf(x7, a, b); }
            if (OP == 2) { x0 = fminf(x0, a) + b; x1 = fminf(x1, a) + b; x2 = fminf(x2, a) + b; x3 = fminf(x3, a) + b; x4 = fminf(x4, a) + b; x5 = fminf(x5, a) + b; x6 = fminf(x6, a) + b; x7 = fminf(x7, a) + b; }
        }
    }
    out[blockIdx.x * 256 + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}
int main() {
    const int blocks = 256 * 16, iters = 256;
    float *out;
    if (hipMalloc(&out, sizeof(float) * blocks * 256) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    const char *names[3] = {"mul+add (2 VALU)", "fma (1 VALU)", "min+add (2 VALU)"};
    const double valuPerChainStep[3] = {2, 1, 2};
    for (int op = 0; op < 3; ++op) {
        for (int rep = 0; rep < 2; ++rep) {
            if (hipEventRecord(e0) != hipSuccess) return 1;
            if (op == 0) hipLaunchKernelGGL(k_peak<0>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 0.001f);
            if (op == 1) hipLaunchKernelGGL(k_peak<1>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 0.001f);
            if (op == 2) hipLaunchKernelGGL(k_peak<2>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 0.001f);
            float ms = 0;
            if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                return 1;
            const double laneOps = double(blocks) * 256 * iters * 8 * 8 * valuPerChainStep[op];
            if (rep == 1) printf("%s: %.3f ms, %.1f T lane-ops/s (VALU instructions x 64 lanes)\n", names[op], ms, laneOps / (ms * 1e-3) / 1e12);
        }
    }
    const hipError_t e = hipGetLastError();
    printf("status %d\n", static_cast<int>(e));
    return e == hipSuccess ? 0 : 1;
}
