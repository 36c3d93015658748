// abi.hip — the C ABI declared in include/restir_abi.h: argument validation, scene-table setup and
// stream-ordered kernel launches. Every entry point returns 0 or a negative RESTIR_E_* code and
// records a message retrievable through restir_last_error().
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.h"
#include "kernels.h"

// ---- static scratch of restir_initial_temporal_noalloc ------------------------------------
// The queued Pass A schedule's scratch (ray queue, random-stream states, shadow bits:
// pass_a_queue_bytes(W*H) bytes) for callers that cannot allocate device memory. Each device holds
// its own copy from module load until the process ends. Host code never uses the name: its
// device address is resolved once per device (rki::resolve_noalloc_scratch) and passed to the
// kernels.
static __device__ __align__(256) unsigned char scratch[RESTIR_NOALLOC_SCRATCH_BYTES];

namespace rki {

int fail(restir_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->error = buf;
    }
    return code;
}

int hip_fail(restir_ctx *ctx, hipError_t e, const char *what) {
    return fail(ctx, RESTIR_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

int use_device(restir_ctx *ctx) {
    // every launching entry point passes here first: bind this context's recorder (or none)
    rk::tl_recorder = ctx->profiling ? &ctx->recorder : nullptr;
    const hipError_t e = hipSetDevice(ctx->device);
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "hipSetDevice");
}

bool overlaps(const void *a, size_t na, const void *b, size_t nb) {
    const char *pa = static_cast<const char *>(a), *pb = static_cast<const char *>(b);
    return pa < pb + nb && pb < pa + na;
}

int check_gbuffer(restir_ctx *ctx, const restir_gbuffer *g, uint32_t W, uint32_t H,
                  bool needDepth, const char *name) {
    if (!g) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: G-buffer is NULL", name);
    if (!g->position || !g->normal || !g->albedo || !g->material || (needDepth && !g->depth))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: G-buffer plane pointer is NULL", name);
    if (g->width != W || g->height != H)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "%s: G-buffer is %ux%u but screenSize is %ux%u",
                    name, g->width, g->height, W, H);
    if (g->plane_stride < static_cast<uint64_t>(W) * H)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "%s: plane_stride %llu < width*height", name,
                    static_cast<unsigned long long>(g->plane_stride));
    if (g->normal_format != RESTIR_NORMAL_F32 && g->normal_format != RESTIR_NORMAL_F16)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: unknown normal format %u", name, g->normal_format);
    if (reinterpret_cast<uintptr_t>(g->normal) % (g->normal_format == RESTIR_NORMAL_F16 ? 2 : 4) != 0)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: normal planes are misaligned", name);
    return RESTIR_OK;
}

rk::GView view_of(const restir_gbuffer *g) {
    rk::GView v{};
    if (!g) return v;
    v.pos = g->position;
    v.nrm = g->normal;
    v.nrmHalf = g->normal_format == RESTIR_NORMAL_F16 ? 1u : 0u;
    v.alb = g->albedo;
    v.mat = g->material;
    v.depth = g->depth;
    v.P = g->plane_stride;
    v.W = g->width;
    v.H = g->height;
    return v;
}

int check_frame(restir_ctx *ctx, const restir_uniforms *u, const char *name) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (!u) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: uniforms are NULL", name);
    if (u->screenSize[0] == 0 || u->screenSize[1] == 0)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: empty screen", name);
    if (static_cast<uint64_t>(u->screenSize[0]) * u->screenSize[1] > (1ull << 31))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: screen too large", name);
    return RESTIR_OK;
}

// Grows the context's scratch to at least `bytes` (hipFree waits for work still using it).
int ensure(restir_ctx *ctx, void **buf, size_t *have, size_t bytes) {
    if (bytes <= *have) return RESTIR_OK;
    if (*buf) {
        const hipError_t e = hipFree(*buf);
        *buf = nullptr;
        *have = 0;
        if (e != hipSuccess) return hip_fail(ctx, e, "scratch: hipFree");
    }
    const hipError_t e = hipMalloc(buf, bytes);
    if (e != hipSuccess) {
        *buf = nullptr;
        return hip_fail(ctx, e, "scratch: hipMalloc");
    }
    *have = bytes;
    return RESTIR_OK;
}

// Workgroups of a persistent launch: as many 256-lane groups as fit a CU (160 KB of LDS at
// `ldsBytes` each, at most 8 groups = 32 waves), on every CU.
uint32_t resident_blocks(uint32_t cus, uint32_t ldsBytes) {
    uint32_t perCu = ldsBytes > 0 ? 163840u / ldsBytes : 8u;
    perCu = perCu > 8u ? 8u : (perCu < 1u ? 1u : perCu);
    return cus * perCu;
}
// the queued tracers' 256-lane groups each hold a `stack`-entry LDS stack per lane and the
// waves' triangle-test lists
uint32_t persistent_blocks(uint32_t cus, int stack) {
    return resident_blocks(cus, static_cast<uint32_t>(stack) * 256u * 4u + rk::kTraceExtraLds);
}

int ensure_scratch(restir_ctx *ctx, size_t bytes) {
    if (bytes > ctx->scratchBytes) ctx->surfValid = false;  // a new allocation holds no records
    return ensure(ctx, &ctx->scratch, &ctx->scratchBytes, bytes);
}

// Binds the device and, for the queued schedule (the default), sizes the context scratch for
// Pass A's ray queue; *queues stays null for the inline schedule. The Pass C ray export is
// invalidated: the queues share the scratch.
int pass_a_queues(restir_ctx *ctx, const restir_launch *launch, uint32_t npx, void **queues) {
    const uint32_t schedule = launch ? launch->schedule : RESTIR_SCHEDULE_QUEUED;
    if (schedule != RESTIR_SCHEDULE_QUEUED && schedule != RESTIR_SCHEDULE_INLINE)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "unknown schedule %u", schedule);
    if (int r = use_device(ctx)) return r;
    *queues = nullptr;
    if (schedule == RESTIR_SCHEDULE_QUEUED) {
        if (int r = ensure_scratch(ctx, rk::pass_a_queue_bytes(npx))) return r;
        *queues = ctx->scratch;
        ctx->queuedNpx = 0;
        ctx->surfValid = false;
    }
    return RESTIR_OK;
}

int pick_tile(restir_ctx *ctx, const restir_launch *l, int *tile) {
    if (!l || (l->tile_x == 0 && l->tile_y == 0)) {
        *tile = 0;
        return RESTIR_OK;
    }
    *tile = rk::tile_index(l->tile_x, l->tile_y);
    if (*tile < 0)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "tile %ux%u is not compiled in", l->tile_x, l->tile_y);
    return RESTIR_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_noallocMu;
void *g_noallocPtr[kMaxDevices] = {};

// The device address of `scratch` on ctx->device, resolved once per device; called by
// restir_create with that device current.
int resolve_noalloc_scratch(restir_ctx *ctx) {
    if (ctx->device >= kMaxDevices)
        return fail(ctx, RESTIR_E_NO_DEVICE, "device %d: at most %d devices", ctx->device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_noallocMu);
    if (!g_noallocPtr[ctx->device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
        g_noallocPtr[ctx->device] = ptr;
    }
    ctx->noallocScratch = g_noallocPtr[ctx->device];
    return RESTIR_OK;
}

}  // namespace rki

using namespace rki;

extern "C" {

int restir_version(void) { return RESTIR_ABI_VERSION; }

int restir_create(restir_ctx **out, int device) {
    if (!out) return RESTIR_E_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RESTIR_E_NO_DEVICE;
    if (device < 0 || device >= n) return RESTIR_E_NO_DEVICE;
    auto *ctx = new restir_ctx();
    ctx->device = device;
    if (use_device(ctx) != RESTIR_OK || hipMalloc(&ctx->counters, rk::kCounterBytes) != hipSuccess ||
        resolve_noalloc_scratch(ctx) != RESTIR_OK) {
        delete ctx;
        return RESTIR_E_HIP;
    }
    ctx->scene.traceHeads = ctx->counters + 64;  // past the 256 bytes of single counters
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->cus = static_cast<uint32_t>(cus);
    ctx->persistentBlocks = persistent_blocks(ctx->cus, ctx->stack);
    *out = ctx;
    return RESTIR_OK;
}

void restir_destroy(restir_ctx *ctx) { delete ctx; }

const char *restir_last_error(const restir_ctx *ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int restir_set_lights(restir_ctx *ctx, const void *pointBuf, size_t pointBytes, const void *triBuf,
                      size_t triBytes, const void *aliasBuf, size_t aliasBytes) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    ctx->lightsSet = false;
    // The clusters were built over the previous light set: their per-light tables are sized for its
    // light count and k_propose reads a snapshot of its alias table, so they must be bound again.
    ctx->clustersSet = false;
    if (int r = use_device(ctx)) return r;
    if (!pointBuf || !aliasBuf)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_lights: point-light and alias buffers are required");
    const size_t hdr = RESTIR_LIGHT_BUFFER_HEADER_BYTES;
    if (pointBytes < hdr || aliasBytes < hdr || (triBuf && triBytes < hdr))
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "set_lights: buffer shorter than its 16-byte header");
    int32_t counts[3] = {0, 0, 0};
    hipError_t e = hipMemcpy(&counts[0], pointBuf, 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && triBuf) e = hipMemcpy(&counts[1], triBuf, 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&counts[2], aliasBuf, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "set_lights: reading buffer headers");
    const int32_t np = counts[0], nt = counts[1], na = counts[2];
    if (np < 0 || nt < 0 || na < 0)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_lights: negative count in a header");
    if (hdr + static_cast<size_t>(np) * sizeof(restir_point_light) > pointBytes)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "set_lights: %d point lights exceed %zu bytes", np, pointBytes);
    if (triBuf && hdr + static_cast<size_t>(nt) * sizeof(restir_tri_light) > triBytes)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "set_lights: %d triangle lights exceed %zu bytes", nt, triBytes);
    if (hdr + static_cast<size_t>(na) * sizeof(restir_alias_table_column) > aliasBytes)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "set_lights: %d alias columns exceed %zu bytes", na, aliasBytes);
    // the alias table covers point lights if there are any, else triangle lights (misc.cpp:426-442),
    // or, with a column for every light of both kinds, the union (point lights first)
    const int32_t expect = np > 0 ? np : nt;
    if (expect == 0) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_lights: no lights");
    const bool mixed = np > 0 && nt > 0 && na == np + nt;
    if (na != expect && !mixed)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "set_lights: alias table has %d columns, expected %d%s", na, expect,
                    np > 0 && nt > 0 ? " (or point + triangle lights for one table over both)" : "");
    const char *pb = static_cast<const char *>(pointBuf);
    const char *ab = static_cast<const char *>(aliasBuf);
    {   // every alias must name a light: the kernels index the light arrays with it
        std::vector<restir_alias_table_column> cols(static_cast<size_t>(na));
        const hipError_t ce = hipMemcpy(cols.data(), ab + hdr, sizeof(restir_alias_table_column) * cols.size(),
                                        hipMemcpyDeviceToHost);
        if (ce != hipSuccess) return hip_fail(ctx, ce, "set_lights: reading alias table");
        for (size_t i = 0; i < cols.size(); ++i)
            if (cols[i].alias < 0 || cols[i].alias >= na)
                return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_lights: alias column %zu names light %d",
                            i, cols[i].alias);
    }
    ctx->scene.points = reinterpret_cast<const restir_point_light *>(pb + hdr);
    ctx->scene.triLights = triBuf ? reinterpret_cast<const restir_tri_light *>(static_cast<const char *>(triBuf) + hdr) : nullptr;
    ctx->scene.alias = reinterpret_cast<const float4 *>(ab + hdr);
    ctx->scene.nPoints = np;
    ctx->scene.nTriLights = nt;
    ctx->scene.nAlias = na;
    ctx->scene.mixed = mixed ? 1 : 0;
    ctx->scene.nLights = mixed ? np + nt : expect;
    ctx->lightsSet = true;
    return RESTIR_OK;
}

int restir_set_bvh(restir_ctx *ctx, const restir_aabb_tree_node *nodes, uint32_t nNodes,
                   const restir_triangle *tris, uint32_t nTris) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    ctx->bvhSet = false;
    if (int r = use_device(ctx)) return r;
    if (!nodes || !tris) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "set_bvh: NULL buffer");
    if (nTris < 2 || nNodes != nTris - 1)
        return fail(ctx, RESTIR_E_BAD_TREE, "set_bvh: %u nodes for %u triangles (need tris-1, tris>=2)", nNodes, nTris);
    std::vector<restir_aabb_tree_node> host(nNodes);
    const hipError_t e = hipMemcpy(host.data(), nodes, sizeof(restir_aabb_tree_node) * nNodes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "set_bvh: reading nodes");
    // validate: every node reached exactly once from node 0, every triangle referenced once
    std::vector<uint8_t> seenNode(nNodes, 0), seenTri(nTris, 0);
    std::vector<std::pair<int32_t, uint32_t>> todo{{0, 1u}};
    uint32_t depth = 0, reached = 0, leaves = 0;
    seenNode[0] = 1;
    while (!todo.empty()) {
        const auto [id, d] = todo.back();
        todo.pop_back();
        ++reached;
        if (d > depth) depth = d;
        for (int32_t c : {host[id].leftChild, host[id].rightChild}) {
            if (c >= 0) {
                if (static_cast<uint32_t>(c) >= nNodes || seenNode[c])
                    return fail(ctx, RESTIR_E_BAD_TREE, "set_bvh: node %d has bad child %d", id, c);
                seenNode[c] = 1;
                todo.push_back({c, d + 1});
            } else {
                const uint32_t t = static_cast<uint32_t>(~c);
                if (t >= nTris || seenTri[t])
                    return fail(ctx, RESTIR_E_BAD_TREE, "set_bvh: node %d has bad leaf %d", id, ~c);
                seenTri[t] = 1;
                ++leaves;
            }
        }
    }
    if (reached != nNodes || leaves != nTris)
        return fail(ctx, RESTIR_E_BAD_TREE, "set_bvh: %u of %u nodes and %u of %u triangles reachable",
                    reached, nNodes, leaves, nTris);
    if (depth > 64)
        return fail(ctx, RESTIR_E_BAD_TREE, "set_bvh: tree depth %u exceeds the 64-entry traversal stack", depth);
    return bind_tree(ctx, nodes, nNodes, tris, depth, nullptr);
}

// Validation shared by the Pass A entry points (check_frame has passed).
int check_pass_a(restir_ctx *ctx, const restir_uniforms *u, const restir_gbuffer *cur, const restir_gbuffer *prev,
                 const restir_reservoir *prevRes, const restir_reservoir *out, const char *name) {
    if (!ctx->lightsSet) return fail(ctx, RESTIR_E_NOT_READY, "%s: lights not set", name);
    const bool vis = (u->flags & RESTIR_VISIBILITY_REUSE_FLAG) != 0;
    const bool temporal = (u->flags & RESTIR_TEMPORAL_REUSE_FLAG) != 0;
    if (vis && !ctx->bvhSet) return fail(ctx, RESTIR_E_NOT_READY, "%s: BVH not set", name);
    const uint32_t W = u->screenSize[0], H = u->screenSize[1];
    if (int r = check_gbuffer(ctx, cur, W, H, false, name)) return r;
    if (!out) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: output is NULL", name);
    const size_t bytes = sizeof(restir_reservoir) * W * H;
    if (temporal) {
        if (int r = check_gbuffer(ctx, prev, W, H, false, name)) return r;
        if (!prevRes) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "%s: previous reservoirs are NULL", name);
        if (overlaps(prevRes, bytes, out, bytes))
            return fail(ctx, RESTIR_E_ALIASING, "%s: output aliases the previous reservoirs", name);
    }
    return RESTIR_OK;
}

int restir_initial_temporal(restir_ctx *ctx, const restir_uniforms *u, const restir_gbuffer *cur,
                            const restir_gbuffer *prev, const restir_reservoir *prevRes,
                            restir_reservoir *out, const restir_launch *launch, void *stream) {
    if (int r = check_frame(ctx, u, "initial_temporal")) return r;
    if (int r = check_pass_a(ctx, u, cur, prev, prevRes, out, "initial_temporal")) return r;
    const uint32_t W = u->screenSize[0], H = u->screenSize[1];
    int tile;
    if (int r = pick_tile(ctx, launch, &tile)) return r;
    void *queues = nullptr;
    if (int r = pass_a_queues(ctx, launch, W * H, &queues)) return r;
    const hipError_t e = rk::launch_pass_a(tile, ctx->stack, *u, ctx->scene, view_of(cur), view_of(prev),
                                           prevRes, out, nullptr, nullptr, queues, ctx->persistentBlocks,
                                           static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "initial_temporal launch");
}

int restir_initial_temporal_noalloc(restir_ctx *ctx, const restir_uniforms *u, const restir_gbuffer *cur,
                                    const restir_gbuffer *prev, const restir_reservoir *prevRes,
                                    restir_reservoir *out, const restir_launch *launch, void *stream) {
    if (int r = check_frame(ctx, u, "initial_temporal_noalloc")) return r;
    const uint32_t W = u->screenSize[0], H = u->screenSize[1];
    const uint32_t schedule = launch ? launch->schedule : RESTIR_SCHEDULE_QUEUED;
    if (schedule != RESTIR_SCHEDULE_QUEUED && schedule != RESTIR_SCHEDULE_INLINE)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "initial_temporal_noalloc: unknown schedule %u", schedule);
    const size_t need = rk::pass_a_queue_bytes(W * H);
    if (schedule == RESTIR_SCHEDULE_QUEUED && need > RESTIR_NOALLOC_SCRATCH_BYTES)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT,
                    "initial_temporal_noalloc: %ux%u pixels need %zu bytes of scratch; the static array "
                    "holds %llu (at most %u pixels)", W, H, need,
                    static_cast<unsigned long long>(RESTIR_NOALLOC_SCRATCH_BYTES), RESTIR_NOALLOC_MAX_PIXELS);
    if (int r = check_pass_a(ctx, u, cur, prev, prevRes, out, "initial_temporal_noalloc")) return r;
    int tile;
    if (int r = pick_tile(ctx, launch, &tile)) return r;
    if (int r = use_device(ctx)) return r;
    if (!ctx->noallocScratch) return fail(ctx, RESTIR_E_HIP, "initial_temporal_noalloc: static scratch unresolved");
    // no per-kernel timing here: the recorder's bookkeeping may allocate host memory
    rk::tl_recorder = nullptr;
    void *queues = schedule == RESTIR_SCHEDULE_QUEUED ? ctx->noallocScratch : nullptr;
    const hipError_t e = rk::launch_pass_a(tile, ctx->stack, *u, ctx->scene, view_of(cur), view_of(prev),
                                           prevRes, out, nullptr, nullptr, queues, ctx->persistentBlocks,
                                           static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "initial_temporal_noalloc launch");
}

int restir_spatial(restir_ctx *ctx, const restir_uniforms *u, int iter, const restir_gbuffer *g,
                   const restir_reservoir *in, restir_reservoir *out, const restir_launch *launch,
                   void *stream) {
    if (int r = check_frame(ctx, u, "spatial")) return r;
    const uint32_t W = u->screenSize[0], H = u->screenSize[1];
    if (int r = check_gbuffer(ctx, g, W, H, true, "spatial")) return r;
    if (!in || !out) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "spatial: reservoir buffer is NULL");
    const size_t bytes = sizeof(restir_reservoir) * W * H;
    if (overlaps(in, bytes, out, bytes)) return fail(ctx, RESTIR_E_ALIASING, "spatial: in and out overlap");
    int tile;
    if (int r = pick_tile(ctx, launch, &tile)) return r;
    if (int r = use_device(ctx)) return r;
    const float cosThr = rt::cos_degrees(u->spatialNormalThreshold);
    const hipError_t e = rk::launch_pass_b(tile, *u, iter, cosThr, view_of(g), in, out,
                                           static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "spatial launch");
}

int restir_unbiased(restir_ctx *ctx, const restir_uniforms *u, int k, int iter,
                    const restir_gbuffer *g, const restir_reservoir *in, restir_reservoir *out,
                    const restir_launch *launch, void *stream) {
    if (int r = check_frame(ctx, u, "unbiased")) return r;
    if (k < 1 || k > 16) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "unbiased: k=%d outside [1,16]", k);
    const bool vis = (u->flags & RESTIR_VISIBILITY_REUSE_FLAG) != 0;
    if (vis && !ctx->bvhSet) return fail(ctx, RESTIR_E_NOT_READY, "unbiased: BVH not set");
    const uint32_t W = u->screenSize[0], H = u->screenSize[1];
    if (int r = check_gbuffer(ctx, g, W, H, false, "unbiased")) return r;
    if (!in || !out) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "unbiased: reservoir buffer is NULL");
    const size_t bytes = sizeof(restir_reservoir) * W * H;
    if (overlaps(in, bytes, out, bytes)) return fail(ctx, RESTIR_E_ALIASING, "unbiased: in and out overlap");
    int tile;
    if (int r = pick_tile(ctx, launch, &tile)) return r;
    uint32_t schedule = launch ? launch->schedule : RESTIR_SCHEDULE_QUEUED;
    if (schedule != RESTIR_SCHEDULE_QUEUED && schedule != RESTIR_SCHEDULE_INLINE)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "unbiased: unknown schedule %u", schedule);
    // a queued neighbour ray carries its pixel in 27 bits of its tag
    if (static_cast<uint64_t>(W) * H >= (1ull << 27)) schedule = RESTIR_SCHEDULE_INLINE;
    if (int r = use_device(ctx)) return r;
    void *queues = nullptr;
    bool pack = true;
    const rk::GView gv = view_of(g);
    if (schedule == RESTIR_SCHEDULE_QUEUED) {
        if (int r = ensure_scratch(ctx, rk::pass_c_queue_bytes(W * H, k))) return r;
        queues = ctx->scratch;
        // RESTIR_INPUT_VISIBLE_FLAG promises the previous Pass C's output over the same G-buffer:
        // when that Pass C ran here on these planes, size and k, its surface records are reused
        const auto &key = ctx->surfKey;
        pack = !((u->flags & RESTIR_INPUT_VISIBLE_FLAG) && ctx->surfValid && key.pos == gv.pos &&
                 key.nrm == gv.nrm && key.P == gv.P && key.W == W && key.H == H && key.nrmHalf == gv.nrmHalf &&
                 key.k == k);
    }
    const hipError_t e = rk::launch_pass_c(tile, ctx->stack, *u, k, iter, ctx->scene, gv, in, out, queues,
                                           ctx->persistentBlocks, static_cast<hipStream_t>(stream), pack);
    ctx->surfValid = queues && e == hipSuccess;
    if (ctx->surfValid) ctx->surfKey = {gv.pos, gv.nrm, gv.P, W, H, gv.nrmHalf, k};
    // the queues hold this call's rays only when it traced any (otherwise they are stale)
    ctx->queuedNpx = (queues && vis && e == hipSuccess) ? W * H : 0;
    ctx->queuedK = k;
    ctx->queuedOut = out;
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "unbiased launch");
}

int restir_shade(restir_ctx *ctx, const float cameraPos[3], float gamma, const restir_gbuffer *g,
                 const restir_reservoir *res, float *rgb, uint64_t rgbStride, void *stream) {
    return restir_shade_ex(ctx, cameraPos, gamma, 0u, g, res, rgb, rgbStride, stream);
}

int restir_shade_ex(restir_ctx *ctx, const float cameraPos[3], float gamma, uint32_t flags, const restir_gbuffer *g,
                    const restir_reservoir *res, float *rgb, uint64_t rgbStride, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (flags & ~RESTIR_SHADE_VISIBILITY_FLAG) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "shade: unknown flags %u", flags);
    const bool vis = (flags & RESTIR_SHADE_VISIBILITY_FLAG) != 0;
    if (!ctx->lightsSet) return fail(ctx, RESTIR_E_NOT_READY, "shade: lights not set");
    if (vis && !ctx->bvhSet) return fail(ctx, RESTIR_E_NOT_READY, "shade: BVH not set (visibility re-test)");
    if (!cameraPos || !res || !rgb) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "shade: NULL argument");
    if (!g) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "shade: G-buffer is NULL");
    if (int r = check_gbuffer(ctx, g, g->width, g->height, false, "shade")) return r;
    if (rgbStride < static_cast<uint64_t>(g->width) * g->height)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "shade: rgb plane stride too small");
    if (!(gamma > 0.0f)) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "shade: gamma must be > 0");
    if (int r = use_device(ctx)) return r;
    const hipError_t e = rk::launch_shade(ctx->scene, rt::f3{cameraPos[0], cameraPos[1], cameraPos[2]},
                                          gamma, view_of(g), res, rgb, rgbStride, vis, ctx->stack,
                                          static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "shade launch");
}

int restir_trace_shadow(restir_ctx *ctx, const float *p1, const float *p2, uint32_t n,
                        uint32_t *shadowed, uint32_t *visits, uint32_t *tests, int loop, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (!ctx->bvhSet) return fail(ctx, RESTIR_E_NOT_READY, "trace_shadow: BVH not set");
    if (n == 0) return RESTIR_OK;
    if (!p1 || !p2 || !shadowed) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "trace_shadow: NULL buffer");
    if ((loop & ~(RESTIR_TRACE_REFERENCE_LOOP | RESTIR_TRACE_ONE_RAY_PER_LANE)) != 0)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "trace_shadow: unknown loop flags %d", loop);
    if (int r = use_device(ctx)) return r;
    const hipError_t e = rk::launch_trace(ctx->stack, (loop & RESTIR_TRACE_REFERENCE_LOOP) != 0,
                                          (loop & RESTIR_TRACE_ONE_RAY_PER_LANE) == 0,
                                          ctx->persistentBlocks, ctx->counters, ctx->scene,
                                          reinterpret_cast<const float4 *>(p1),
                                          reinterpret_cast<const float4 *>(p2), n, shadowed, visits,
                                          tests, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "trace_shadow launch");
}

int restir_ground_truth(restir_ctx *ctx, const float cameraPos[3], const restir_gbuffer *g,
                        const uint32_t *pixels, uint32_t n, uint32_t nTriSamples, double *rgb,
                        uint64_t rgbStride, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (!ctx->lightsSet || !ctx->bvhSet) return fail(ctx, RESTIR_E_NOT_READY, "ground_truth: lights or BVH not set");
    if (n == 0) return RESTIR_OK;
    if (!cameraPos || !pixels || !rgb || !g)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "ground_truth: NULL argument");
    if (int r = check_gbuffer(ctx, g, g->width, g->height, false, "ground_truth")) return r;
    if (rgbStride < n) return fail(ctx, RESTIR_E_SIZE_MISMATCH, "ground_truth: rgb stride %llu < n",
                                   static_cast<unsigned long long>(rgbStride));
    if (n > (1u << 24)) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "ground_truth: n=%u > 2^24", n);
    if (ctx->scene.nTriLights > 0 && (nTriSamples == 0 || nTriSamples > 65536))
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "ground_truth: n_tri_samples=%u not in [1, 65536]",
                    nTriSamples);
    uint32_t sq = 1;
    while (sq * sq < nTriSamples) ++sq;
    if (int r = use_device(ctx)) return r;
    // the pixel list is device memory: launched lanes index the G-buffer with it, so it is
    // checked on the host first (one small copy)
    std::vector<uint32_t> host(n);
    hipError_t e = hipMemcpy(host.data(), pixels, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "ground_truth: reading the pixel list");
    const uint64_t npx = static_cast<uint64_t>(g->width) * g->height;
    for (uint32_t i = 0; i < n; ++i)
        if (host[i] >= npx) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "ground_truth: pixel %u out of range", host[i]);
    if (int r = ensure(ctx, &ctx->truthScratch, &ctx->truthScratchBytes, rk::truth_scratch_bytes(ctx->scene, n)))
        return r;
    e = rk::launch_truth(ctx->stack, ctx->scene, view_of(g), cameraPos, pixels, n, sq, ctx->truthScratch, rgb,
                         rgbStride, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "ground_truth launch");
}

int restir_primary_gbuffer(restir_ctx *ctx, const restir_camera *cam, const int32_t *triMaterial,
                           const float *materials, uint32_t nMaterials, float *position,
                           void *normal, uint32_t normalFormat, float *albedo, float *material, float *depth,
                           uint64_t planeStride, uint32_t width, uint32_t height, void *stream) {
    if (!ctx) return RESTIR_E_INVALID_ARGUMENT;
    if (!ctx->bvhSet) return fail(ctx, RESTIR_E_NOT_READY, "primary_gbuffer: BVH not set");
    if (!cam || !triMaterial || !materials || !position || !normal || !albedo || !material || !depth)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "primary_gbuffer: NULL argument");
    if (normalFormat != RESTIR_NORMAL_F32 && normalFormat != RESTIR_NORMAL_F16)
        return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "primary_gbuffer: unknown normal format %u", normalFormat);
    if (nMaterials == 0) return fail(ctx, RESTIR_E_INVALID_ARGUMENT, "primary_gbuffer: no materials");
    if (width == 0 || height == 0 || planeStride < static_cast<uint64_t>(width) * height)
        return fail(ctx, RESTIR_E_SIZE_MISMATCH, "primary_gbuffer: bad size");
    if (int r = use_device(ctx)) return r;
    const rk::CamBasis cb = rk::make_cam_basis(*cam, static_cast<float>(width) / static_cast<float>(height));
    const hipError_t e = rk::launch_primary(ctx->stack, ctx->scene, cb, triMaterial, materials, position,
                                            normal, normalFormat == RESTIR_NORMAL_F16, albedo, material, depth,
                                            planeStride, width, height,
                                            static_cast<hipStream_t>(stream));
    return e == hipSuccess ? RESTIR_OK : hip_fail(ctx, e, "primary_gbuffer launch");
}

int restir_camera_projection_view(const restir_camera *cam, float aspect, float out[16]) {
    if (!cam || !out || !(aspect > 0.0f)) return RESTIR_E_INVALID_ARGUMENT;
    const rk::CamBasis b = rk::make_cam_basis(*cam, aspect);
    const float f = 1.0f / tanf(0.5f * cam->fovYRadians);
    const float view[4][4] = {{b.right.x, b.right.y, b.right.z, -rt::dot(b.right, b.pos)},
                              {b.down.x, b.down.y, b.down.z, -rt::dot(b.down, b.pos)},
                              {b.fwd.x, b.fwd.y, b.fwd.z, -rt::dot(b.fwd, b.pos)},
                              {0.0f, 0.0f, 0.0f, 1.0f}};
    const float proj[4][4] = {{f / aspect, 0.0f, 0.0f, 0.0f},
                              {0.0f, f, 0.0f, 0.0f},
                              {0.0f, 0.0f, b.A, b.B},
                              {0.0f, 0.0f, 1.0f, 0.0f}};
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            float acc = 0.0f;
            for (int k = 0; k < 4; ++k) acc += proj[r][k] * view[k][c];
            out[c * 4 + r] = acc;
        }
    return RESTIR_OK;
}

}  // extern "C"
