// kernels.h — launch interface between the C ABI (abi.hip) and the gfx950 kernels (passes.hip).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/restir_abi.h"
#include "rt_traverse.h"

namespace rk {

// ---- optional per-kernel timing (restir_profile_enable). Every launch goes through RK_LAUNCH,
// which hands the kernel a start/stop event pair when the calling thread has a recorder bound;
// the events bracket exactly that kernel on its stream (hipExtLaunchKernel).
struct Recorder;
extern thread_local Recorder *tl_recorder;
void rec_pair(const char *name, hipStream_t stream, hipEvent_t *start, hipEvent_t *stop);

#define RK_LAUNCH(name, kernel, grid, block, shm, stream, ...)                                   \
    do {                                                                                         \
        hipEvent_t rk_start_ = nullptr, rk_stop_ = nullptr;                                      \
        if (::rk::tl_recorder) ::rk::rec_pair(name, stream, &rk_start_, &rk_stop_);              \
        hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, rk_start_, rk_stop_, 0u,         \
                              __VA_ARGS__);                                                      \
    } while (0)

// Scene tables as the kernels see them: arrays past the 16-byte SSBO headers, counts read once.
struct SceneView {
    const restir_point_light *points;
    const restir_tri_light *triLights;
    const float4 *alias;  // aliasTableColumn as float4 {prob, alias bits, oriProb, aliasOriProb}
    int32_t nPoints, nTriLights, nAlias;
    int32_t mixed;    // 1: the alias table covers point lights then triangle lights (the union)
    int32_t nLights;  // lights the alias table indexes: nAlias
    const rt::Node *nodes;
    const float4 *tris;     // 3 float4 per triangle
    const rt::WNode *wide;  // the 4-entry nodes over `nodes` (same indices), built by the library
    int32_t treeNested;     // 1: every box finite and inside the box stored for its parent
    float maxCoord;         // largest |box coordinate| of the tree
    uint32_t *traceHeads;   // the queue tracer's work counters (the context's): kTraceShards words,
                            // kTraceShardStride apart, cleared by each tracer launch
};
// The queue tracer's rays are taken in chunks of kTraceChunk through kTraceShards counters, each in
// a 128-byte line of its own: device-wide atomics on one address are served one at a time (about
// 11 ns each on the MI355X), on different lines in parallel.
constexpr uint32_t kTraceShards = 64, kTraceShardStride = 32, kTraceChunk = 64;
constexpr size_t kCounterBytes = 256 + sizeof(uint32_t) * kTraceShards * kTraceShardStride;

struct GView {
    const float *pos;
    const void *nrm;  // 3 float planes, or 3 half planes when nrmHalf
    const float *alb, *mat, *depth;
    uint64_t P;  // plane stride in elements
    uint32_t W, H;
    uint32_t nrmHalf;
};

// The normal of pixel i, from float or half planes (half -> float is exact).
__device__ __forceinline__ rt::f3 load_normal(const GView &g, size_t i) {
    if (g.nrmHalf) {
        const _Float16 *h = static_cast<const _Float16 *>(g.nrm);
        return rt::f3{static_cast<float>(h[i]), static_cast<float>(h[g.P + i]), static_cast<float>(h[2 * g.P + i])};
    }
    const float *f = static_cast<const float *>(g.nrm);
    return rt::f3{f[i], f[g.P + i], f[2 * g.P + i]};
}

struct Tile {
    uint32_t tx, ty;
};

// Tile shapes compiled in (pixels per workgroup = tx*ty). The first is the default.
constexpr int kNumTiles = 5;
extern const Tile kTiles[kNumTiles];
int tile_index(uint32_t tx, uint32_t ty);  // -1 if not compiled

// candLight/candQ: the learned proposal's candidates ([M][npx]), or null for alias-table RIS.
// queueScratch == nullptr: one dispatch, the visibility ray traced inline per pixel. Otherwise
// the ray-queue schedule, using pass_a_queue_bytes(npx) bytes of device scratch.
// LDS of a queued tracer's 256-lane workgroup besides its `stack`-entry stack per lane: each
// wave's list of waiting triangle tests (fewer than 64 left after a round plus at most 4 per lane
// from one visit; a 4-byte triangle index and an owner byte each) and a tally word per lane.
constexpr uint32_t kTraceListEntries = 64 + 4 * 64;
constexpr uint32_t kTraceExtraLds = 4u * kTraceListEntries * 5u + 256u * 4u;
hipError_t launch_pass_a(int tile, int stack, const restir_uniforms &u, const SceneView &sc,
                         const GView &cur, const GView &prev, const restir_reservoir *prevRes,
                         restir_reservoir *out, const int32_t *candLight, const float *candQ,
                         void *queueScratch, uint32_t gridCap, hipStream_t s);
size_t pass_a_queue_bytes(uint32_t npx);
hipError_t launch_pass_b(int tile, const restir_uniforms &u, int iter, float cosThr,
                         const GView &g, const restir_reservoir *in, restir_reservoir *out,
                         hipStream_t s);
// queueScratch == nullptr: one dispatch, rays traced inline per pixel. Otherwise the ray-queue
// schedule, using pass_c_queue_bytes(npx, k) bytes of device scratch.
hipError_t launch_pass_c(int tile, int stack, const restir_uniforms &u, int k, int iter,
                         const SceneView &sc, const GView &g, const restir_reservoir *in,
                         restir_reservoir *out, void *queueScratch, uint32_t gridCap, hipStream_t s,
                         bool packSurface = true);
size_t pass_c_queue_bytes(uint32_t npx, int k);
// Copies the shadow rays that the last queued Pass C over (npx, k) put in `queueScratch`
// (which = 0: centre rays; 1: the neighbour rays of pixels whose centre ray found the light; the
// light end of both is read from that pass's output `merged`); *n receives the number of rays.
hipError_t pass_c_rays(const void *queueScratch, uint32_t npx, int k, const restir_reservoir *merged, int which,
                       float4 *p1, float4 *p2, uint32_t cap, uint32_t *n, hipStream_t s);
// visibility: the final visibility re-test (tree traversal with a `stack`-entry stack)
hipError_t launch_shade(const SceneView &sc, rt::f3 cam, float gamma, const GView &g,
                        const restir_reservoir *res, float *rgb, uint64_t rgbP, bool visibility, int stack,
                        hipStream_t s);
// persistent: fixed grid of at most gridCap workgroups pulling rays through *counter (which the
// launch clears with a kernel, so the sequence can be captured in a graph).
hipError_t launch_trace(int stack, bool refLoop, bool persistent, uint32_t gridCap,
                        uint32_t *counter, const SceneView &sc, const float4 *p1, const float4 *p2,
                        uint32_t n, uint32_t *shadowed, uint32_t *visits, uint32_t *tests,
                        hipStream_t s);

// clears n words (a kernel rather than hipMemsetAsync, so the sequence can be graph-captured)
__global__ void k_zero_u32(uint32_t *p, uint32_t n);
// clears n < 64 words of p and the queue tracer's counters (one workgroup of 64 lanes)
__global__ void k_zero_counters(uint32_t *p, uint32_t n, uint32_t *heads);

// ---- batched views: one persistent launch over the tiles of n views (each view its own
// uniforms, scene tables, G-buffers and reservoirs); `table` holds views_table_bytes(n) of
// device memory, `counter` one word.
struct ViewA {
    restir_uniforms u;
    SceneView sc;
    GView cur, prev;
    const restir_reservoir *prevRes;
    restir_reservoir *out;
};
struct ViewB {
    restir_uniforms u;
    GView g;
    const restir_reservoir *in;
    restir_reservoir *out;
    float cosThr;
};
struct ViewC {
    restir_uniforms u;
    SceneView sc;
    GView g;
    const restir_reservoir *in;
    restir_reservoir *out;
};
struct ViewShade {
    SceneView sc;
    rt::f3 cam;
    GView g;
    const restir_reservoir *res;
    float *rgb;
    uint64_t rgbP;
};
size_t views_table_bytes(uint32_t n);
hipError_t launch_pass_a_views(int tile, int stack, const ViewA *views, uint32_t n, void *table,
                               uint32_t *counter, uint32_t gridCap, hipStream_t s);
hipError_t launch_pass_b_views(int tile, const ViewB *views, uint32_t n, int iter, void *table,
                               uint32_t *counter, uint32_t gridCap, hipStream_t s);
hipError_t launch_pass_c_views(int tile, int stack, const ViewC *views, uint32_t n, int k, int iter,
                               void *table, uint32_t *counter, uint32_t gridCap, hipStream_t s);
hipError_t launch_shade_views(const ViewShade *views, uint32_t n, float gamma, void *table,
                              uint32_t *counter, uint32_t gridCap, hipStream_t s);

// ---- learned proposal (mlp.hip)
size_t mlp_scratch_bytes(uint32_t n, int K);
hipError_t mlp_init(const restir_mlp &m, uint64_t seed, const float *prior, hipStream_t s);
hipError_t mlp_features(const GView &g, rt::f3 cam, const float lo[3], const float hi[3],
                        const uint32_t *pixels, uint32_t n, uint16_t *out, hipStream_t s);
hipError_t mlp_forward(const restir_mlp &m, const uint16_t *X, uint32_t n, float *logits,
                       void *scratch, hipStream_t s);
hipError_t mlp_loss_grad(const restir_mlp &m, const uint16_t *X, const int32_t *target,
                         const uint32_t *mask, uint32_t n, float *gw, float *gb, float *loss,
                         void *scratch, hipStream_t s);
hipError_t mlp_train(restir_mlp &m, const GView &g, rt::f3 cam, const float lo[3], const float hi[3],
                     const restir_reservoir *res, const int32_t *lightCluster, int32_t nLights, int32_t triBase,
                     uint32_t n, float lr,
                     uint32_t seed, float *lossOut, void *scratch, hipStream_t s);
hipError_t mlp_adam(restir_mlp &m, const float *gw, const float *gb, const uint32_t *nValid, float lr,
                    hipStream_t s);
hipError_t mlp_propose(const restir_mlp &m, const GView &g, rt::f3 cam, const float lo[3],
                       const float hi[3], const SceneView &sc, const int32_t *lightCluster,
                       const float *lightPWithin, const int32_t *clusterBegin,
                       const int32_t *clusterCount, const int32_t *clusterLights,
                       const restir_alias_table_column *clusterAlias, const float4 *clusterFused, uint32_t M,
                       float mix, uint32_t seed, int32_t *candLight, float *candQ, hipStream_t s);

// ---- FP64 ground truth (truth.hip)
constexpr uint32_t kTruthChunk = 128;  // lights per workgroup row
size_t truth_scratch_bytes(const SceneView &sc, uint32_t n);
hipError_t launch_truth(int stack, const SceneView &sc, const GView &g, const float cam[3],
                        const uint32_t *pixels, uint32_t n, uint32_t sq, void *scratch, double *rgb,
                        uint64_t stride, hipStream_t s);

struct CamBasis {
    rt::f3 pos, fwd, right, down;
    float sx, sy, A, B;
};
CamBasis make_cam_basis(const restir_camera &c, float aspect);
hipError_t launch_primary(int stack, const SceneView &sc, const CamBasis &cb,
                          const int32_t *triMaterial, const float *materials, float *pos,
                          void *nrm, bool nrmHalf, float *alb, float *mat, float *depth, uint64_t P, uint32_t W,
                          uint32_t H, hipStream_t s);

}  // namespace rk
