// ctx.h — the state behind a restir_ctx handle (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.h"

namespace rk {
// Event pairs of the kernels launched while profiling is on, in launch order.
struct Recorder {
    struct Span {
        const char *name;  // a string literal at the launch site
        hipEvent_t start, stop;
    };
    std::vector<Span> spans;
    std::vector<hipEvent_t> pool;  // recycled events
    hipEvent_t take();
    void clear();
    ~Recorder();
};
}  // namespace rk

struct restir_ctx {
    int device = 0;
    std::string error;
    rk::SceneView scene{};
    bool lightsSet = false;
    bool bvhSet = false;
    int stack = 32;
    uint32_t treeDepth = 0;
    uint32_t *counters = nullptr;  // device work counters of the persistent kernels
    uint32_t cus = 256;                // compute units of the device
    uint32_t persistentBlocks = 1280;  // workgroups of the persistent tracers (see persistent_blocks)
    void *noallocScratch = nullptr;  // this device's copy of abi.hip's static `scratch` array
    void *scratch = nullptr;       // ray queues of the queued Pass C schedule
    size_t scratchBytes = 0;
    uint32_t queuedNpx = 0;        // shape of the last queued Pass C (its rays stay in scratch)
    const restir_reservoir *queuedOut = nullptr;  // its output (the light end of its neighbour rays)
    int queuedK = 0;
    // the G-buffer planes, size and neighbour count whose surface records (k_pack_surface) the
    // last queued Pass C left in scratch; surfValid is cleared when anything else writes scratch
    struct {
        const void *pos, *nrm;
        uint64_t P;
        uint32_t W, H, nrmHalf;
        int k;
    } surfKey{};
    bool surfValid = false;
    bool profiling = false;
    rk::Recorder recorder;
    void *mlpScratch = nullptr;    // training tensors of the learned proposal
    size_t mlpScratchBytes = 0;
    void *truthScratch = nullptr;  // per-light-chunk partial sums of the FP64 ground truth
    size_t truthScratchBytes = 0;
    void *viewTable = nullptr;     // per-view records of the batched-view launches
    size_t viewTableBytes = 0;
    void *bvhScratch = nullptr;    // primitive order, keys and ranges of the GPU tree build
    size_t bvhScratchBytes = 0;
    void *bvhInfoHost = nullptr;   // pinned: each tree level's node and split counts
    void *aliasScratch = nullptr;  // powers, weights and queues of the device alias-table build
    size_t aliasScratchBytes = 0;
    void *wide = nullptr;          // the bound tree's 4-entry nodes (+ the nesting flag)
    size_t wideBytes = 0;
    // light clusters of the learned proposal (device pointers, caller-owned)
    bool clustersSet = false;
    int32_t nClusters = 0;
    const int32_t *lightCluster = nullptr;
    const float *lightPWithin = nullptr;
    const int32_t *clusterBegin = nullptr, *clusterCount = nullptr, *clusterLights = nullptr;
    const restir_alias_table_column *clusterAlias = nullptr;
    // per cluster slot, built from the tables above: {prob, light of the column, light of its
    // alias, p_within of the column}, {p_within of the alias, p_src of both lights, 0}
    void *clusterFused = nullptr;
    size_t clusterFusedBytes = 0;
    ~restir_ctx() {
        if (counters) (void)hipFree(counters);
        if (scratch) (void)hipFree(scratch);
        if (mlpScratch) (void)hipFree(mlpScratch);
        if (truthScratch) (void)hipFree(truthScratch);
        if (viewTable) (void)hipFree(viewTable);
        if (bvhScratch) (void)hipFree(bvhScratch);
        if (bvhInfoHost) (void)hipHostFree(bvhInfoHost);
        if (aliasScratch) (void)hipFree(aliasScratch);
        if (wide) (void)hipFree(wide);
        if (clusterFused) (void)hipFree(clusterFused);
    }
};

// helpers shared by the ABI files (defined in abi.hip; bind_tree in bvh_build.hip)
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
int pass_a_queues(restir_ctx *ctx, const restir_launch *launch, uint32_t npx, void **queues);
int ensure(restir_ctx *ctx, void **buf, size_t *have, size_t bytes);
uint32_t resident_blocks(uint32_t cus, uint32_t ldsBytes);
uint32_t persistent_blocks(uint32_t cus, int stack);
int bind_tree(restir_ctx *ctx, const void *nodes, uint32_t nNodes, const void *tris, uint32_t depth, hipStream_t s);
}  // namespace rki
