// cluster_build.hip — the k-means light clustering of the learned proposal with its nearest-
// centroid step on the GPU (no reference counterpart). Everything else (initial centroids,
// centroid means in double in light order, grouping, per-cluster Vose tables) is the host build's
// own code (host_build.cpp), so the two builds return the same bytes.
//
// The step is n x k distances per Lloyd iteration: 1M lights x 4,096 clusters is 4.3 G distances,
// minutes on one host core and about a millisecond here. Each lane takes kPerLane lights; the
// workgroup stages the centroids through LDS in chunks of kChunk (one 16-byte broadcast read per
// centroid serves the lane's kPerLane lights). Distance and tie rule are the host's:
// d = (dx*dx + dy*dy) + dz*dz in fp32 without contraction, the first minimum wins.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "clusters.h"
#include "ctx.h"

using namespace rki;

namespace {

constexpr uint32_t kBlock = 256, kPerLane = 4, kChunk = 1024;

__global__ __launch_bounds__(kBlock) void k_assign_nearest(const float4 *__restrict__ pos, uint32_t n,
                                                           const float4 *__restrict__ cent, uint32_t k,
                                                           uint32_t *__restrict__ assign) {
    __shared__ float4 c[kChunk];
    const uint32_t base = blockIdx.x * kBlock * kPerLane + threadIdx.x;
    float x[kPerLane], y[kPerLane], z[kPerLane], bestD[kPerLane];
    uint32_t best[kPerLane];
#pragma unroll
    for (uint32_t i = 0; i < kPerLane; ++i) {
        const uint32_t l = base + i * kBlock;
        const float4 p = l < n ? pos[l] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        x[i] = p.x;
        y[i] = p.y;
        z[i] = p.z;
        bestD[i] = 0.0f;
        best[i] = 0u;
    }
    for (uint32_t j0 = 0; j0 < k; j0 += kChunk) {
        const uint32_t m = k - j0 < kChunk ? k - j0 : kChunk;
        __syncthreads();  // the previous chunk is no longer read
        for (uint32_t t = threadIdx.x; t < m; t += kBlock) c[t] = cent[j0 + t];
        __syncthreads();
        for (uint32_t t = 0; t < m; ++t) {
            const float4 q = c[t];
            const uint32_t j = j0 + t;
#pragma unroll
            for (uint32_t i = 0; i < kPerLane; ++i) {
                const float dx = x[i] - q.x, dy = y[i] - q.y, dz = z[i] - q.z;
                const float d = (dx * dx + dy * dy) + dz * dz;
                if (j == 0 || d < bestD[i]) {
                    bestD[i] = d;
                    best[i] = j;
                }
            }
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < kPerLane; ++i) {
        const uint32_t l = base + i * kBlock;
        if (l < n) assign[l] = best[i];
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int restir_build_light_clusters_device(restir_ctx *ctx, const restir_point_light *points, uint32_t n,
                                                  uint32_t k, uint64_t seed, uint32_t iterations,
                                                  int32_t *lightCluster, float *lightPWithin, int32_t *clusterBegin,
                                                  int32_t *clusterCount, int32_t *clusterLights,
                                                  restir_alias_table_column *clusterAlias, float *centroidsOut) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (!points || k == 0 || k > n)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "build_light_clusters_device: n=%u k=%u", n, k);
    if (int r = use_device(ctx)) return r;
    std::vector<float4> pos(n);
    for (uint32_t l = 0; l < n; ++l) pos[l] = make_float4(points[l].pos[0], points[l].pos[1], points[l].pos[2], 0.0f);
    DevBuf dPos, dCent, dAssign;
    hipError_t e = hipMalloc(&dPos.p, sizeof(float4) * n);
    if (e == hipSuccess) e = hipMalloc(&dCent.p, sizeof(float4) * k);
    if (e == hipSuccess) e = hipMalloc(&dAssign.p, sizeof(uint32_t) * n);
    if (e == hipSuccess) e = hipMemcpy(dPos.p, pos.data(), sizeof(float4) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(ctx, e, "build_light_clusters_device: buffers");
    std::vector<float4> cent(k);
    const NearestStep device = [&](const float *c, uint32_t *assign) {
        for (uint32_t j = 0; j < k; ++j) cent[j] = make_float4(c[3 * j], c[3 * j + 1], c[3 * j + 2], 0.0f);
        hipError_t err = hipMemcpy(dCent.p, cent.data(), sizeof(float4) * k, hipMemcpyHostToDevice);
        if (err != hipSuccess) return hip_fail(ctx, err, "build_light_clusters_device: centroids");
        const uint32_t blocks = (n + kBlock * kPerLane - 1) / (kBlock * kPerLane);
        hipLaunchKernelGGL(k_assign_nearest, dim3(blocks), dim3(kBlock), 0, nullptr,
                           static_cast<const float4 *>(dPos.p), n, static_cast<const float4 *>(dCent.p), k,
                           static_cast<uint32_t *>(dAssign.p));
        err = hipGetLastError();
        if (err == hipSuccess) err = hipMemcpy(assign, dAssign.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
        return err == hipSuccess ? 0 : hip_fail(ctx, err, "build_light_clusters_device: nearest step");
    };
    const int r = build_light_clusters(points, n, k, seed, iterations, device, lightCluster, lightPWithin,
                                       clusterBegin, clusterCount, clusterLights, clusterAlias, centroidsOut);
    if (r == RESTIR_E_INVALID_ARGUMENT) return fail(ctx, r, "build_light_clusters_device: NULL output");
    return r;
}
