// ctx.h — the state behind a restir_ctx handle (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "kernels.h"

struct restir_ctx {
    int device = 0;
    std::string error;
    rk::SceneView scene{};
    bool lightsSet = false;
    bool bvhSet = false;
    int stack = 32;
    uint32_t treeDepth = 0;
    uint32_t *counters = nullptr;  // device work counters of the persistent kernels
    uint32_t persistentBlocks = 1280;
    void *scratch = nullptr;       // ray queues of the queued Pass C schedule
    size_t scratchBytes = 0;
    void *mlpScratch = nullptr;    // training tensors of the learned proposal
    size_t mlpScratchBytes = 0;
    // light clusters of the learned proposal (device pointers, caller-owned)
    bool clustersSet = false;
    int32_t nClusters = 0;
    const int32_t *lightCluster = nullptr;
    const float *lightPWithin = nullptr;
    const int32_t *clusterBegin = nullptr, *clusterCount = nullptr, *clusterLights = nullptr;
    const restir_alias_table_column *clusterAlias = nullptr;
    ~restir_ctx() {
        if (counters) (void)hipFree(counters);
        if (scratch) (void)hipFree(scratch);
        if (mlpScratch) (void)hipFree(mlpScratch);
    }
};

// helpers shared by abi.hip and abi_mlp.hip (defined in abi.hip)
namespace rki {
int fail(restir_ctx *ctx, int code, const char *fmt, ...);
int hip_fail(restir_ctx *ctx, hipError_t e, const char *what);
int use_device(restir_ctx *ctx);
bool overlaps(const void *a, size_t na, const void *b, size_t nb);
int check_gbuffer(restir_ctx *ctx, const restir_gbuffer *g, uint32_t W, uint32_t H, bool needDepth,
                  const char *name);
rk::GView view_of(const restir_gbuffer *g);
int check_frame(restir_ctx *ctx, const restir_uniforms *u, const char *name);
int pick_tile(restir_ctx *ctx, const restir_launch *l, int *tile);
int ensure(restir_ctx *ctx, void **buf, size_t *have, size_t bytes);
}  // namespace rki
