/*
 * restir_abi.h — C ABI of the MI355X (gfx950) ReSTIR hot path.
 *
 * The entry points stand in for the compute-shader dispatches of lukedan/ReSTIR-Vulkan:
 *
 *   restir_initial_temporal  <- RestirPass::issueCommands software branch
 *                               (src/passes/restirPass.h:36-53, shader src/shaders/restirOmni.glsl:86-212)
 *   restir_spatial           <- SpatialReusePass::issueCommands
 *                               (src/passes/spatialReusePass.h:11-26, shader src/shaders/spatialReuse.comp:30-85)
 *   restir_unbiased          <- UnbiasedReusePass::issueCommands software branch
 *                               (src/passes/unbiasedReusePass.h:23-40, shader src/shaders/unbiasedReuse.glsl:50-185)
 *   restir_trace_shadow      <- testVisibility/raytrace (src/shaders/include/visibilityTest.glsl:1-30,
 *                               src/shaders/include/softwareRaytracing.glsl:42-85) as a batch over rays
 *   restir_shade             <- LightingPass GBUFFER_DEBUG_NONE branch (src/shaders/lighting.frag:49-71, 103)
 *   restir_build_aabb_tree   <- AabbTree::build (src/aabbTreeBuilder.cpp:52-214)
 *   restir_build_alias_table <- createAliasTable (src/misc.cpp:418-497)
 *   restir_set_lights        <- descriptor set 0 bindings 0-2 of RestirPass (restirPass.h:371-423),
 *                               buffers packed as SceneBuffers::create does (src/sceneBuffers.h:100-124, 241-270)
 *   restir_set_bvh           <- descriptor set 2 of RestirPass / set 1 of UnbiasedReusePass
 *                               (AabbTreeBuffers::create, src/aabbTreeBuilder.h:25-51)
 *
 * Conventions: every call returns 0 on success or a negative RESTIR_E_* code; restir_last_error()
 * returns the message of the last failure on that context. Device pointers are caller-owned.
 * Work is ordered on the caller's HIP stream (passed as an opaque pointer; NULL = default stream).
 * A pass never reads the buffer it writes (the reference always ping-pongs; in == out is rejected).
 * One context per host thread.
 */
#ifndef RESTIR_ABI_H
#define RESTIR_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------------------------------
 * Wire formats. Byte-identical to the reference's GLSL std430/std140 structs.
 * -------------------------------------------------------------------------------------------- */

/* LightSample, src/shaders/include/structs/restirStructs.glsl:19-29 (48 B). */
typedef struct restir_light_sample {
    float position_emissionLum[4]; /* xyz light sample position, w emission luminance */
    float normal[4];               /* xyz light normal, w > 0.5: triangle-light normal valid */
    int32_t lightIndex;            /* >= 0 point light; < 0 triangle light -1-idx */
    float pHat;
    float sumWeights;
    float w;
} restir_light_sample;

/* Reservoir with RESERVOIR_SIZE 1, restirStructs.glsl:31-34 (64 B, 16-byte aligned). */
typedef struct restir_reservoir {
    restir_light_sample samples[1];
    uint32_t numStreamSamples;
    uint32_t _pad[3];
} restir_reservoir;

/* RestirUniforms, restirStructs.glsl:40-56 (std140, 128 B). Matrix is column-major. */
typedef struct restir_uniforms {
    float prevFrameProjectionViewMatrix[16];
    float cameraPos[4];
    uint32_t screenSize[2];
    uint32_t frame;
    uint32_t initialLightSampleCount;
    uint32_t temporalSampleCountMultiplier;
    float spatialPosThreshold;
    float spatialNormalThreshold; /* degrees */
    uint32_t spatialNeighbors;
    float spatialRadius;
    int32_t flags;
    uint32_t _pad[2];
} restir_uniforms;

#define RESTIR_VISIBILITY_REUSE_FLAG (1 << 0) /* restirStructs.glsl:37 */
#define RESTIR_TEMPORAL_REUSE_FLAG (1 << 1)   /* restirStructs.glsl:38 */
/* No counterpart in the reference (its default behaviour is this flag off): the estimator
 * without the reference's two departures from RIS. Pass A ends its candidate loop with the
 * standard contribution weight W = sumW / (M p_hat) instead of addSampleToReservoir's
 * selection-time W (reservoir.glsl:28-42), and a triangle-light candidate's density is
 * p / area (uniform by area) instead of p / (|cos| area) (restirOmni.glsl:129-131), whose
 * estimator converges to the integral of f |cos_light| rather than of f. */
#define RESTIR_STANDARD_WEIGHTS_FLAG (1 << 2)
/* No counterpart in the reference: a promise about a Pass C's input, read by restir_unbiased only.
 * Every input reservoir with W > 0 holds a sample whose segment from its own pixel was found
 * unoccluded, as the output of a restir_unbiased with RESTIR_VISIBILITY_REUSE_FLAG over the same
 * G-buffer guarantees (its centre re-test leaves W = 0 otherwise): the second and later Pass C of a
 * frame may set it. The pass then does not trace the rays that retrace such a segment: the
 * centre ray when the merged sample is the pixel's own, and the neighbour rays from the pixel
 * whose reservoir supplied the merged sample. On the queued schedule, when the previous call on
 * this context was a queued restir_unbiased over the same G-buffer planes, size and k (no queued
 * Pass A since), the pass also reuses the position and normal records that call packed from the
 * planes, which the promise implies are unchanged. The results are the same as without the flag
 * when the promise holds; when it does not, those rays count as unoccluded. */
#define RESTIR_INPUT_VISIBLE_FLAG (1 << 3)

/* AabbTreeNode, src/shaders/include/structs/aabbTree.glsl:1-8 (80 B). Child < 0 is leaf ~triIndex. */
typedef struct restir_aabb_tree_node {
    float leftAabbMin[4];
    float leftAabbMax[4];
    float rightAabbMin[4];
    float rightAabbMax[4];
    int32_t leftChild;
    int32_t rightChild;
    int32_t _pad[2];
} restir_aabb_tree_node;

/* Triangle, aabbTree.glsl:9-13 (48 B, w = 1). */
typedef struct restir_triangle {
    float p1[4];
    float p2[4];
    float p3[4];
} restir_triangle;

/* pointLight, src/shaders/include/structs/light.glsl:2-5 (32 B). */
typedef struct restir_point_light {
    float pos[4];
    float color_luminance[4]; /* w = luminance */
} restir_point_light;

/* triLight, light.glsl:7-13 (80 B). */
typedef struct restir_tri_light {
    float p1[4];
    float p2[4];
    float p3[4];
    float emission_luminance[4]; /* w = luminance */
    float normalArea[4];         /* xyz unit normal, w area */
} restir_tri_light;

/* aliasTableColumn, light.glsl:15-20 (16 B). */
typedef struct restir_alias_table_column {
    float prob;
    int32_t alias;
    float oriProb;
    float aliasOriProb;
} restir_alias_table_column;

/* Light/alias buffers are the reference's SSBO images: a 16-byte header whose first int is the
 * element count, then the array (sceneBuffers.h:100-124, misc.h:28-30). */
#define RESTIR_LIGHT_BUFFER_HEADER_BYTES 16

/* G-buffer as structure-of-arrays float planes, row-major, row pitch = width. Replaces the
 * sampler2D images of the reference (gBufferPass.cpp:75-108, gBuffer.frag:27-80).
 * position/normal: 3 planes (x, y, z). albedo: 4 planes (r, g, b, emissive flag > 0.5).
 * material: 2 planes (roughness, metallic). depth: 1 plane of hardware depth (z/w, [0,1]).
 * A pixel whose normal is (0,0,0) is background. plane_stride is in elements (>= width*height).
 * Normals may be stored as IEEE half planes (RESTIR_NORMAL_F16; the 4K stress workload's format):
 * every pass reads them exactly as their float values. */
typedef struct restir_gbuffer {
    const float *position;
    const void *normal;      /* 3 planes of float, or of IEEE half (uint16) with RESTIR_NORMAL_F16 */
    const float *albedo;
    const float *material;
    const float *depth;
    uint64_t plane_stride;   /* in elements of each plane's type */
    uint32_t width;
    uint32_t height;
    uint32_t normal_format;  /* RESTIR_NORMAL_F32 (0) or RESTIR_NORMAL_F16 */
    uint32_t reserved;       /* 0 */
} restir_gbuffer;
#define RESTIR_NORMAL_F32 0u
#define RESTIR_NORMAL_F16 1u

/* Launch shape of the per-pixel passes: the workgroup tile (pixels per workgroup = tile_x *
 * tile_y, a multiple of 64; {0,0} = built-in default) and the schedule of the shadow rays of
 * Pass A (visibility reuse) and Pass C: RESTIR_SCHEDULE_QUEUED (default) moves them into queues
 * traced by a persistent kernel, using context-owned scratch; RESTIR_SCHEDULE_INLINE traces them
 * per pixel in a single dispatch, as the reference does. Both produce identical reservoirs.
 * (Pass C falls back to the inline schedule at 2^27 pixels and more.) */
#define RESTIR_SCHEDULE_QUEUED 0
#define RESTIR_SCHEDULE_INLINE 1
typedef struct restir_launch {
    uint32_t tile_x;
    uint32_t tile_y;
    uint32_t schedule;
    uint32_t _reserved;
} restir_launch;

/* ----------------------------------------------------------------------------------------------
 * Error codes
 * -------------------------------------------------------------------------------------------- */
#define RESTIR_OK 0
#define RESTIR_E_INVALID_ARGUMENT (-1)
#define RESTIR_E_SIZE_MISMATCH (-2)
#define RESTIR_E_NOT_READY (-3)   /* lights or BVH not set */
#define RESTIR_E_HIP (-4)         /* HIP runtime failure */
#define RESTIR_E_BAD_TREE (-5)    /* child index out of range, node count != tris - 1, too deep */
#define RESTIR_E_ALIASING (-6)    /* output buffer aliases an input buffer */
#define RESTIR_E_NO_DEVICE (-7)

typedef struct restir_ctx restir_ctx;

/* ----------------------------------------------------------------------------------------------
 * Context
 * -------------------------------------------------------------------------------------------- */
int restir_create(restir_ctx **out, int device);
void restir_destroy(restir_ctx *ctx);
const char *restir_last_error(const restir_ctx *ctx);
/* ABI revision; bumped whenever an entry point's signature or documented behaviour changes. Bindings compare it with the
 * revision they were written against and refuse a mismatched library. */
#define RESTIR_ABI_VERSION 17
int restir_version(void);

/* Light tables: device pointers to the reference buffer images (header + array). The counts are
 * read back once here and validated against the byte sizes. A NULL triangle buffer means none.
 * Replaces descriptor set 0 bindings 0..2 of restirOmni.glsl:11-23. As in the reference, the
 * alias table covers the point lights if there are any, else the triangle lights (its column
 * count must match). Extension: with both kinds present, an alias table of n_point + n_tri
 * columns (restir_build_alias_table_mixed) is one table over the point lights followed by the
 * triangle lights, and Pass A samples both kinds from it (a triangle light is drawn as the
 * reference draws it, restirOmni.glsl:122-131). Binding lights unbinds the light clusters
 * (restir_set_light_clusters): they describe the previous light set, so the learned-proposal
 * entry points return RESTIR_E_NOT_READY until clusters are bound again. */
int restir_set_lights(restir_ctx *ctx, const void *point_lights, size_t point_bytes,
                      const void *tri_lights, size_t tri_bytes, const void *alias_table,
                      size_t alias_bytes);

/* BVH: device pointers in the AabbTreeBuffers layout (aabbTreeBuilder.h:25-51). The tree is
 * read back once to validate child indices and to measure its depth. */
int restir_set_bvh(restir_ctx *ctx, const restir_aabb_tree_node *nodes, uint32_t n_nodes,
                   const restir_triangle *tris, uint32_t n_tris);

/* AabbTree::build (aabbTreeBuilder.cpp:52-214) on the device, in place of the host build plus
 * restir_set_bvh for scenes whose geometry changes: the tree restir_build_aabb_tree makes, node
 * for node and bit for bit, from device vertex data (n_tris x 9 floats: p1, p2, p3), written to
 * device buffers nodes[n_tris - 1] and tris[n_tris] and bound to ctx as restir_set_bvh binds them
 * (the build's own bookkeeping replaces set_bvh's read-back check). depth_out (host, may be NULL)
 * gets the tree depth. The build waits for the stream once per tree level (it reads each level's
 * node count back), so it cannot be captured into a graph. Context scratch: about 380 bytes per
 * triangle. A non-finite coordinate gives RESTIR_E_INVALID_ARGUMENT. */
int restir_build_bvh(restir_ctx *ctx, const float *tri_xyz, uint32_t n_tris,
                     restir_aabb_tree_node *nodes, restir_triangle *tris, uint32_t *depth_out,
                     void *stream);

/* Pass A: initial RIS + visibility reuse + temporal reuse (restirOmni.glsl:86-212).
 * prev / prev_res are read only when RESTIR_TEMPORAL_REUSE_FLAG is set. */
int restir_initial_temporal(restir_ctx *ctx, const restir_uniforms *u, const restir_gbuffer *cur,
                            const restir_gbuffer *prev, const restir_reservoir *prev_res,
                            restir_reservoir *out, const restir_launch *launch, void *stream);

/* No counterpart in the reference: restir_initial_temporal for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results (bit for bit); the queued schedule's scratch (ray queue,
 * random-stream states, shadow bits: 44 bytes per pixel + 1 KB) comes from a static device array
 * of RESTIR_NOALLOC_SCRATCH_BYTES that each device holds from library load until the process ends,
 * instead of the context's allocation. Calls no hipMalloc/hipFree; frames above
 * RESTIR_NOALLOC_MAX_PIXELS pixels are rejected (RESTIR_E_INVALID_ARGUMENT) on the queued schedule
 * (the inline schedule needs no scratch). All contexts on a device share its array, which keeps
 * the last call's contents: calls that overlap on one device (other streams, other threads) race.
 * Per-kernel timing (restir_profile_enable) does not cover this entry point. */
#define RESTIR_NOALLOC_SCRATCH_BYTES (96ull << 20)
#define RESTIR_NOALLOC_MAX_PIXELS 2287792u /* e.g. 1920x1080 or 2048x1080 */
int restir_initial_temporal_noalloc(restir_ctx *ctx, const restir_uniforms *u,
                                    const restir_gbuffer *cur, const restir_gbuffer *prev,
                                    const restir_reservoir *prev_res, restir_reservoir *out,
                                    const restir_launch *launch, void *stream);

/* Pass B: biased spatial reuse, push constant iter (spatialReuse.comp:30-85). */
int restir_spatial(restir_ctx *ctx, const restir_uniforms *u, int iter, const restir_gbuffer *g,
                   const restir_reservoir *in, restir_reservoir *out, const restir_launch *launch,
                   void *stream);

/* Pass C: unbiased spatial reuse with k neighbours, 1 <= k <= 16 (unbiasedReuse.glsl:50-185,
 * where k is the compile-time NUM_NEIGHBORS = 3). The pass seeds its generator with
 * frame*17 + iter; iter = 0 is the reference's single pass, iter > 0 decorrelates further passes. */
int restir_unbiased(restir_ctx *ctx, const restir_uniforms *u, int k, int iter,
                    const restir_gbuffer *g, const restir_reservoir *in, restir_reservoir *out,
                    const restir_launch *launch, void *stream);

/* Final shading (lighting.frag:49-71, 103): rgb = sum p_hat_full(y) * W, emissive pixels show
 * their albedo, background pixels are 0, then pow(rgb, 1/gamma). Writes 3 float planes. */
int restir_shade(restir_ctx *ctx, const float camera_pos[3], float gamma, const restir_gbuffer *g,
                 const restir_reservoir *res, float *rgb_planes, uint64_t rgb_plane_stride,
                 void *stream);
/* restir_shade with options (no counterpart in the reference's lighting pass). With
 * RESTIR_SHADE_VISIBILITY_FLAG, the final shading re-tests visibility: a pixel whose sample has
 * nonzero W traces the segment from its surface to the sample (testVisibility semantics,
 * visibilityTest.glsl:1-30) and shades to 0 when it is shadowed (the BVH must be set). After an
 * unbiased Pass C this changes nothing (a shadowed centre already has W = 0); after Pass B it
 * removes the light the biased reuse carries into shadow. */
#define RESTIR_SHADE_VISIBILITY_FLAG 1u
int restir_shade_ex(restir_ctx *ctx, const float camera_pos[3], float gamma, uint32_t flags,
                    const restir_gbuffer *g, const restir_reservoir *res, float *rgb_planes,
                    uint64_t rgb_plane_stride, void *stream);

/* Shadow rays p1 -> p2 (float4 each, w ignored) through testVisibility semantics: the segment is
 * shrunk by 0.001 at both ends. shadowed[i] = 1 if occluded. node_visits / tri_tests may be NULL.
 * loop = RESTIR_TRACE_REFERENCE_LOOP reproduces the reference's loop exactly (triangle hits take
 * effect every 8 node visits, softwareRaytracing.glsl:39-77), so node_visits equals its
 * loop-iteration count; RESTIR_TRACE_FIRST_HIT stops a ray at its first hit. Both give the same
 * shadowed[] values. By default a persistent grid pulls rays from a queue and refills idle lanes
 * of a wave; RESTIR_TRACE_ONE_RAY_PER_LANE (OR-ed into `loop`) launches one lane per ray instead.
 * The context's work counter is shared, so traces on one context are stream-ordered. */
#define RESTIR_TRACE_FIRST_HIT 0
#define RESTIR_TRACE_REFERENCE_LOOP 1
#define RESTIR_TRACE_ONE_RAY_PER_LANE 2
int restir_trace_shadow(restir_ctx *ctx, const float *p1, const float *p2, uint32_t n,
                        uint32_t *shadowed, uint32_t *node_visits, uint32_t *tri_tests, int loop,
                        void *stream);

/* ----------------------------------------------------------------------------------------------
 * Host-side builders (no device needed)
 * -------------------------------------------------------------------------------------------- */

/* AabbTree::build over n_tris triangles given in scene order (tri_xyz: 9 floats per triangle).
 * Writes n_tris - 1 nodes and the n_tris triangles (w = 1). n_tris must be >= 2. Returns the tree
 * depth (root = 1) through *depth_out when non-NULL. */
int restir_build_aabb_tree(const float *tri_xyz, uint32_t n_tris, restir_aabb_tree_node *nodes,
                           restir_triangle *tris, uint32_t *depth_out);

/* createAliasTable: point lights if n_point > 0, else triangle lights. Writes n columns. */
int restir_build_alias_table(const restir_point_light *points, uint32_t n_point,
                             const restir_tri_light *tris, uint32_t n_tri,
                             restir_alias_table_column *columns);

/* createAliasTable (misc.cpp:418-497) on the device, for lights whose power changes: the columns
 * restir_build_alias_table makes, bit for bit, from device light records (point lights if
 * n_point > 0, else triangle lights; plain arrays, no SSBO header) into device columns[n]. One
 * workgroup, one launch, no host round trip (it may be captured into a graph); the work is
 * ordered on `stream`. Context scratch: 20 bytes per light. */
int restir_build_alias_table_device(restir_ctx *ctx, const restir_point_light *points,
                                    uint32_t n_point, const restir_tri_light *tris, uint32_t n_tri,
                                    restir_alias_table_column *columns, void *stream);

/* One alias table over n_point point lights followed by n_tri triangle lights (no reference
 * counterpart: createAliasTable covers one kind): weights luminance and luminance * area, built by
 * createAliasTable's FIFO construction (misc.cpp:418-497). Writes n_point + n_tri columns. */
int restir_build_alias_table_mixed(const restir_point_light *points, uint32_t n_point,
                                   const restir_tri_light *tris, uint32_t n_tri,
                                   restir_alias_table_column *columns);

/* Light clusters for the learned proposal (no reference counterpart): seeded k-means over the n
 * point-light positions into k clusters (1 <= k <= n; initial centroids by a PCG32(seed)
 * partial Fisher-Yates draw, `iterations` Lloyd steps, ties to the lower centroid), then
 * per-cluster Vose tables over luminance as createAliasTable builds them (misc.cpp:418-497).
 * Outputs: light_cluster[n], light_pwithin[n] (luminance share within its cluster),
 * cluster_begin[k], cluster_count[k], cluster_lights[n] (lights grouped by cluster, index
 * order), cluster_alias[n] (alias entries local to the cluster), centroids[3k] (may be NULL).
 * The clusters must cover the alias table's index space: with one table over both kinds, pass
 * after the point lights one record per triangle light with its centroid as position and
 * luminance * area as luminance. */
int restir_build_light_clusters(const restir_point_light *points, uint32_t n, uint32_t k,
                                uint64_t seed, uint32_t iterations, int32_t *light_cluster,
                                float *light_pwithin, int32_t *cluster_begin,
                                int32_t *cluster_count, int32_t *cluster_lights,
                                restir_alias_table_column *cluster_alias, float *centroids);

/* restir_build_light_clusters with its nearest-centroid step (n x k distances per Lloyd
 * iteration) on ctx's device; the rest of the build is the host build's own code, so the outputs
 * are byte-identical to restir_build_light_clusters for the same arguments. Same host inputs and
 * outputs; synchronous; allocates and frees its device buffers (16 B per light and per cluster,
 * 4 B per light). For light counts where the host step takes minutes (1M lights into 4,096
 * clusters). No reference counterpart. */
int restir_build_light_clusters_device(restir_ctx *ctx, const restir_point_light *points, uint32_t n,
                                       uint32_t k, uint64_t seed, uint32_t iterations,
                                       int32_t *light_cluster, float *light_pwithin,
                                       int32_t *cluster_begin, int32_t *cluster_count,
                                       int32_t *cluster_lights, restir_alias_table_column *cluster_alias,
                                       float *centroids);

/* ----------------------------------------------------------------------------------------------
 * Learned light proposal (no reference counterpart). A 128-128-128-128-K ReLU MLP over per-pixel
 * features (position in the scene box, normal, view direction; raw plus 6 sin/cos octaves = 114,
 * zero-padded to 128) gives a softmax pi_x over K light clusters. Candidates are drawn from
 *   q(l | x) = (1 - mix) * pi_x(cluster(l)) * p_within(l) + mix * p_src(l)
 * (p_src: the reference's alias-table distribution), so RIS with weights p_hat/q stays unbiased;
 * mix = 0 is rejected. Weights are bf16 with an fp32 master copy; all products run on MFMA with
 * fp32 accumulation. K is 64, 128 or 256 (a quad's K numerators held in LDS), or a multiple of
 * 128 up to 4096 (the output layer streamed through LDS in 128-cluster blocks; at most 8
 * candidates per pixel). All pointers in restir_mlp are device pointers.
 * -------------------------------------------------------------------------------------------- */
typedef struct restir_mlp {
    uint32_t n_clusters;    /* K */
    uint32_t adam_step;     /* Adam steps taken; restir_mlp_train increments it */
    uint16_t *weights_bf16; /* W1, W2, W3 (128x128) then W4 (Kx128), each [out][in] */
    float *weights;         /* fp32 master copy, same layout */
    float *bias;            /* b1, b2, b3 (128 each) then b4 (K) */
    float *adam_m;          /* first moments: weights then biases */
    float *adam_v;          /* second moments: weights then biases */
    uint32_t *adam_step_device; /* optional device copy of the step count (1 uint32), or NULL.
                                 * When set, restir_mlp_train increments it on the device and the
                                 * Adam step reads it there, so a sequence of training steps can be
                                 * captured into a HIP graph and replayed; adam_step (host) then
                                 * only counts the calls. restir_mlp_init zeroes it. */
} restir_mlp;

#define RESTIR_MLP_FEATURES 128
/* element counts of the parameter arrays for K clusters */
#define RESTIR_MLP_N_WEIGHTS(K) (3u * 128u * 128u + (uint32_t)(K) * 128u)
#define RESTIR_MLP_N_BIAS(K) (3u * 128u + (uint32_t)(K))

/* Seeded initialisation: He-uniform weights, zero biases except the output bias = out_prior
 * (device array of K floats, e.g. log cluster power shares; -30000 masks an empty cluster),
 * zeroed Adam moments, adam_step = 0 (and *adam_step_device = 0 when set). */
int restir_mlp_init(restir_ctx *ctx, restir_mlp *mlp, uint64_t seed, const float *out_prior,
                    void *stream);

/* Binds the cluster tables of restir_build_light_clusters (device copies) for the proposal.
 * Validates them (copying them to the host, synchronously) and builds from them a context-owned
 * device table of 32 bytes per light (each cluster slot's light, p_within and p_src, for its
 * column and its alias); the caller's buffers must stay alive while the context uses them.
 * Call it after restir_set_lights: the slot table snapshots the bound alias table, and binding
 * new lights unbinds the clusters. */
int restir_set_light_clusters(restir_ctx *ctx, uint32_t k, const int32_t *light_cluster,
                              const float *light_pwithin, const int32_t *cluster_begin,
                              const int32_t *cluster_count, const int32_t *cluster_lights,
                              const restir_alias_table_column *cluster_alias);

/* Draws M <= 16 candidates per pixel from q: cand_light[j * W*H + p], cand_q[j * W*H + p]
 * (light index, or -1 with q = 0 for an empty cluster). Background pixels are left untouched.
 * The MLP is evaluated once per 2x2 pixel quad, at the quad's first surface pixel in the order
 * (0,0), (1,0), (0,1), (1,1); every pixel of the quad draws from that distribution with its own
 * random stream, and cand_q is the density it drew from. scene_lo/hi bound the positions fed
 * to the features. Seeded by (u->frame, pixel). */
int restir_learned_candidates(restir_ctx *ctx, const restir_uniforms *u, const restir_gbuffer *g,
                              const restir_mlp *mlp, const float scene_lo[3],
                              const float scene_hi[3], uint32_t M, float mix, int32_t *cand_light,
                              float *cand_q, void *stream);

/* Pass A with RIS over given candidates instead of the alias table (u->initialLightSampleCount
 * of them per pixel); visibility and temporal reuse as restir_initial_temporal. */
int restir_initial_temporal_candidates(restir_ctx *ctx, const restir_uniforms *u,
                                       const restir_gbuffer *cur, const restir_gbuffer *prev,
                                       const restir_reservoir *prev_res, restir_reservoir *out,
                                       const int32_t *cand_light, const float *cand_q,
                                       const restir_launch *launch, void *stream);

/* One online training step: n_samples (a multiple of 64, <= 65536) pseudo-random pixels of the
 * frame whose final reservoir holds a light with W > 0 become (features, cluster of that light)
 * pairs; mean cross-entropy of the softmax, backward pass, Adam (beta 0.9/0.999, eps 1e-8) at
 * learning rate lr. *loss_out (device float, may be NULL) receives the mean loss. */
int restir_mlp_train(restir_ctx *ctx, restir_mlp *mlp, const restir_uniforms *u,
                     const restir_gbuffer *g, const restir_reservoir *res,
                     const float scene_lo[3], const float scene_hi[3], uint32_t n_samples,
                     float lr, uint32_t seed, float *loss_out, void *stream);

/* Building blocks exposed for checks: features of listed pixels (bf16, n x 128); logits of the
 * MLP (fp32, n x K) for bf16 features (n a multiple of 64, features and weights_bf16 16-byte
 * aligned); loss and gradients (fp32, parameter layouts) for features, targets and a 0/1 mask. */
int restir_mlp_features(restir_ctx *ctx, const restir_gbuffer *g, const float camera_pos[3],
                        const float scene_lo[3], const float scene_hi[3], const uint32_t *pixels,
                        uint32_t n, uint16_t *features, void *stream);
int restir_mlp_forward(restir_ctx *ctx, const restir_mlp *mlp, const uint16_t *features,
                       uint32_t n, float *logits, void *stream);
int restir_mlp_loss_grad(restir_ctx *ctx, const restir_mlp *mlp, const uint16_t *features,
                         const int32_t *target, const uint32_t *mask, uint32_t n, float *grad_w,
                         float *grad_b, float *loss, void *stream);
/* The Adam step of restir_mlp_train alone: from gradients grad_w, grad_b (fp32, parameter
 * layouts) it advances adam_step (and *adam_step_device when set) and updates the fp32 master
 * weights, biases, moments and the bf16 weight copy (beta 0.9 / 0.999, eps 1e-8, rate lr). */
int restir_mlp_adam(restir_ctx *ctx, restir_mlp *mlp, const float *grad_w, const float *grad_b,
                    float lr, void *stream);

/* ----------------------------------------------------------------------------------------------
 * Scene-side helpers for synthetic inputs (no reference counterpart; the reference rasterises).
 * -------------------------------------------------------------------------------------------- */

/* Primary-ray G-buffer over the BVH set on ctx (closest hit, ties to the lower triangle index).
 * Camera follows src/camera.h:25-50: view rows (right, -up, forward), projection with
 * f = 1/tan(fovY/2), depth = (zFar*z - zNear*zFar)/((zFar - zNear)*z).
 * tri_material: one material id per triangle; materials: 8 floats each
 * (albedo r,g,b, emissive flag, roughness, metallic, 0, 0). Writes all G-buffer planes. */
typedef struct restir_camera {
    float position[3];
    float lookAt[3];
    float worldUp[3];
    float fovYRadians;
    float zNear;
    float zFar;
} restir_camera;

int restir_primary_gbuffer(restir_ctx *ctx, const restir_camera *cam, const int32_t *tri_material,
                           const float *materials, uint32_t n_materials, float *position,
                           void *normal, uint32_t normal_format, float *albedo, float *material,
                           float *depth, uint64_t plane_stride, uint32_t width, uint32_t height,
                           void *stream);

/* ---- the reference's G-buffer attachments (gBufferPass.cpp:75-108, written by gBuffer.frag:27-80)
 * as copied into device buffers (vkCmdCopyImageToBuffer, or images in exported memory with linear
 * tiling): row y of an attachment starts at data + y * row_bytes, texels packed in a row. */
typedef enum restir_texel_format {
    RESTIR_TEXEL_RGBA8_SRGB = 1,   /* VK_FORMAT_R8G8B8A8_SRGB: albedo (rgb decoded sRGB -> linear, a unorm) */
    RESTIR_TEXEL_RGBA8_UNORM = 2,  /* VK_FORMAT_R8G8B8A8_UNORM */
    RESTIR_TEXEL_RGB16_SNORM = 3,  /* VK_FORMAT_R16G16B16_SNORM: normal, the reference's first choice */
    RESTIR_TEXEL_RGBA16_SNORM = 4, /* VK_FORMAT_R16G16B16A16_SNORM */
    RESTIR_TEXEL_RGB16_SFLOAT = 5, /* VK_FORMAT_R16G16B16_SFLOAT */
    RESTIR_TEXEL_RGBA16_SFLOAT = 6,/* VK_FORMAT_R16G16B16A16_SFLOAT */
    RESTIR_TEXEL_RGB32_SFLOAT = 7, /* VK_FORMAT_R32G32B32_SFLOAT: world position, normal fallback */
    RESTIR_TEXEL_RGBA32_SFLOAT = 8,/* VK_FORMAT_R32G32B32A32_SFLOAT */
    RESTIR_TEXEL_RG16_UNORM = 9,   /* VK_FORMAT_R16G16_UNORM: roughness, metallic */
    RESTIR_TEXEL_D32_SFLOAT = 10,  /* depth aspect of VK_FORMAT_D32_SFLOAT or D32_SFLOAT_S8_UINT */
    RESTIR_TEXEL_D24_UNORM = 11    /* depth aspect of VK_FORMAT_D24_UNORM_S8_UINT: 32-bit texels,
                                      depth in the low 24 bits */
} restir_texel_format;

typedef struct restir_attachment {
    const void *data;   /* device pointer */
    uint32_t format;    /* restir_texel_format */
    uint32_t row_bytes; /* >= width * texel size */
} restir_attachment;

typedef struct restir_gbuffer_attachments {
    restir_attachment albedo;          /* RGBA8_SRGB, RGBA8_UNORM or RGBA32_SFLOAT; a > 0.5: emissive */
    restir_attachment normal;          /* RGB16/RGBA16 SNORM or SFLOAT, RGB32/RGBA32 SFLOAT */
    restir_attachment material;        /* RG16_UNORM or RGBA32_SFLOAT (roughness, metallic) */
    restir_attachment world_position;  /* RGB32_SFLOAT or RGBA32_SFLOAT */
    restir_attachment depth;           /* D32_SFLOAT or D24_UNORM */
    uint32_t width, height;
} restir_gbuffer_attachments;

/* No counterpart in the reference, which samples its attachments: decodes them into the G-buffer
 * planes restir_gbuffer describes (the arguments as restir_primary_gbuffer's outputs), the way a
 * Vulkan sampler reads them: SNORM16 c -> max(c / 32767, -1), UNORM16 c / 65535, UNORM8 c / 255,
 * sRGB8 rgb through the sRGB EOTF (c/255 <= 0.04045 ? (c/255) / 12.92 : ((c/255 + 0.055) / 1.055)^2.4,
 * evaluated in double and rounded), half -> float exactly, D24 c / (2^24 - 1); normals are written
 * as float or half planes (RESTIR_NORMAL_*). One pass over the pixels, stream-ordered. */
int restir_gbuffer_from_attachments(restir_ctx *ctx, const restir_gbuffer_attachments *att,
                                    float *position, void *normal, uint32_t normal_format,
                                    float *albedo, float *material, float *depth,
                                    uint64_t plane_stride, void *stream);

/* The 4x4 projection-view matrix of camera.h:25-50 for the given aspect, column-major, as the
 * reference writes it into RestirUniforms::prevFrameProjectionViewMatrix (app.cpp:779). */
int restir_camera_projection_view(const restir_camera *cam, float aspect, float out_col_major[16]);

/* ---- batched views (no counterpart in the reference, which renders one view): one persistent
 * launch per pass covers the tiles of n views, each with its own uniforms, G-buffers,
 * reservoirs and scene (the lights and BVH set on `scene`, a context on the same device). The
 * pixels are computed by the same code as the single-view entry points, so each view's output is
 * bit-identical to calling restir_initial_temporal / restir_spatial / restir_unbiased (inline
 * schedule) / restir_shade for it alone. Workgroups take tiles from a global counter in view
 * order, so small and large views share the GPU without a tail per view. The per-view records
 * are written to context-owned device memory by kernels on `stream` (no host-memory reads at
 * execution), so a sequence of these calls can be captured in a HIP graph; calls on one context
 * must stay on one stream. Outputs may not overlap each other or any view's inputs. */
typedef struct restir_view {
    const restir_ctx *scene;         /* lights + BVH of this view */
    const restir_uniforms *uniforms; /* screenSize, frame, flags, ... of this view */
    const restir_gbuffer *cur;       /* G-buffer of this frame */
    const restir_gbuffer *prev;      /* previous frame's (Pass A with temporal reuse) */
    const restir_reservoir *in;      /* Pass A: previous reservoirs; B, C: input; shade: final */
    restir_reservoir *out;           /* Pass A, B, C output */
    float *rgb;                      /* restir_shade_views output planes */
    uint64_t rgb_stride;
} restir_view;

#define RESTIR_MAX_VIEWS 64
int restir_initial_temporal_views(restir_ctx *ctx, const restir_view *views, uint32_t n,
                                  const restir_launch *launch, void *stream);
int restir_spatial_views(restir_ctx *ctx, const restir_view *views, uint32_t n, int iter,
                         const restir_launch *launch, void *stream);
int restir_unbiased_views(restir_ctx *ctx, const restir_view *views, uint32_t n, int k, int iter,
                          const restir_launch *launch, void *stream);
int restir_shade_views(restir_ctx *ctx, const restir_view *views, uint32_t n, float gamma,
                       void *stream);

/* FP64 ground truth of the direct lighting at `n` listed pixels (device array of pixel indices
 * y*width+x): for each, the sum over every light of Le * BRDF * G * V in double precision - the
 * integral that the passes estimate (integrand restirUtils.glsl:30-55, visibility of the segment
 * shrunk by 0.001 at both ends as visibilityTest.glsl:1-30). Triangle lights are integrated with
 * n_tri_samples stratified samples each (rounded up to a square). Emissive pixels return their
 * emission, background pixels 0. Writes rgb[c * rgb_stride + i] (device doubles). No counterpart
 * in the reference: it is the reference of the error metric of converged images. */
int restir_ground_truth(restir_ctx *ctx, const float cameraPos[3], const restir_gbuffer *g,
                        const uint32_t *pixels, uint32_t n, uint32_t n_tri_samples, double *rgb,
                        uint64_t rgb_stride, void *stream);

/* ---- measurement (no counterpart in the reference, which times nothing; for benchmarks and
 * profilers of a host that calls the entry points above) */

/* Per-kernel timing. While enabled, every kernel this context launches is bracketed by a pair of
 * HIP events on its stream (hipExtLaunchKernel start/stop events, so the span is exactly the
 * kernel). restir_profile_read waits for the recorded kernels, writes one row per kernel name
 * (at most `capacity` rows; *count receives the number of names) and clears the record. */
typedef struct restir_kernel_time {
    char name[48];
    uint32_t launches;
    float total_ms;
    float max_ms;
    uint32_t _reserved;
} restir_kernel_time;  /* 64 B */
int restir_profile_enable(restir_ctx *ctx, int enable);
int restir_profile_read(restir_ctx *ctx, restir_kernel_time *out, uint32_t capacity,
                        uint32_t *count);

/* The shadow rays of the last queued-schedule restir_unbiased call on this context, unless a
 * queued Pass A ran since (the queues share scratch) (which = 0: the centre re-tests; 1: the
 * neighbour tests of pixels whose centre ray found the light), as p1/p2 float4 pairs in queue
 * order (p1.w holds a pixel tag, ignored by restir_trace_shadow). The light end of a ray is read
 * from that call's output reservoirs: export before they change. Copies min(number of rays, capacity) rays on `stream`
 * and stores the number of rays in *count (a host pointer; the call synchronizes `stream`). Lets a
 * benchmark count node visits and triangle tests of exactly the rays the traversal kernel traced. */
int restir_unbiased_rays(restir_ctx *ctx, int which, float *p1, float *p2, uint32_t capacity,
                         uint32_t *count, void *stream);
/* Kept for callers of ABI 14, where it switched on the copies of the centre rays that
 * restir_unbiased_rays(which = 0) returns. The queued Pass C now keeps them in any case (its tracer
 * reads them from the queue), so the switch has no effect. */
int restir_set_ray_export(restir_ctx *ctx, int on);

#ifdef __cplusplus
}

static_assert(sizeof(restir_light_sample) == 48, "LightSample is 48 B");
static_assert(sizeof(restir_reservoir) == 64, "Reservoir is 64 B");
static_assert(offsetof(restir_reservoir, numStreamSamples) == 48, "numStreamSamples @48");
static_assert(offsetof(restir_light_sample, lightIndex) == 32, "lightIndex @32");
static_assert(offsetof(restir_light_sample, w) == 44, "w @44");
static_assert(sizeof(restir_uniforms) == 128, "RestirUniforms is 128 B");
static_assert(offsetof(restir_uniforms, cameraPos) == 64, "cameraPos @64");
static_assert(offsetof(restir_uniforms, screenSize) == 80, "screenSize @80");
static_assert(offsetof(restir_uniforms, frame) == 88, "frame @88");
static_assert(offsetof(restir_uniforms, initialLightSampleCount) == 92, "M @92");
static_assert(offsetof(restir_uniforms, temporalSampleCountMultiplier) == 96, "mult @96");
static_assert(offsetof(restir_uniforms, spatialPosThreshold) == 100, "posThr @100");
static_assert(offsetof(restir_uniforms, spatialNormalThreshold) == 104, "norThr @104");
static_assert(offsetof(restir_uniforms, spatialNeighbors) == 108, "k @108");
static_assert(offsetof(restir_uniforms, spatialRadius) == 112, "radius @112");
static_assert(offsetof(restir_uniforms, flags) == 116, "flags @116");
static_assert(sizeof(restir_aabb_tree_node) == 80, "AabbTreeNode is 80 B");
static_assert(offsetof(restir_aabb_tree_node, leftChild) == 64, "leftChild @64");
static_assert(offsetof(restir_aabb_tree_node, rightChild) == 68, "rightChild @68");
static_assert(sizeof(restir_triangle) == 48, "Triangle is 48 B");
static_assert(sizeof(restir_point_light) == 32, "pointLight is 32 B");
static_assert(sizeof(restir_tri_light) == 80, "triLight is 80 B");
static_assert(sizeof(restir_alias_table_column) == 16, "aliasTableColumn is 16 B");
#endif

#endif /* RESTIR_ABI_H */
