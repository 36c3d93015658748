// alias_build.hip — createAliasTable (src/misc.cpp:418-497) on the device: the reference's table,
// column for column and bit for bit, in one workgroup and one launch (no host round trip, so it
// can sit inside a captured graph). The kernel body is alias_kernel.h, shared with a CPU build
// (one thread per lane) that runs it under ThreadSanitizer and AddressSanitizer.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ctx.h"
#include "kernels.h"

#define AK_FN __device__ __forceinline__
#include "alias_kernel.h"

using namespace rki;

namespace {

constexpr uint32_t kLanes = ak::kLanes;

// the workgroup policy of alias_kernel.h on the device
struct DeviceGroup {
    __device__ __forceinline__ uint32_t tid() const { return threadIdx.x; }
    __device__ __forceinline__ void sync() const { __syncthreads(); }
    __device__ __forceinline__ uint32_t shfl_up(uint32_t v, uint32_t d) const {
        return __shfl_up(v, static_cast<unsigned int>(d), 64);
    }
};

__global__ __launch_bounds__(kLanes) void k_alias_build(const restir_point_light *points, const restir_tri_light *triLights,
                                                        uint32_t n, restir_alias_table_column *cols, float *power,
                                                        float *scaled, uint32_t *under, uint32_t *overA,
                                                        uint32_t *overB) {
    __shared__ ak::Shared sh;
    DeviceGroup g;
    ak::build(g, sh, points, triLights, n, cols, power, scaled, under, overA, overB);
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
