/*
 * restir_oracle.h — CPU oracle for the ReSTIR hot path. TEST INFRASTRUCTURE ONLY.
 *
 * This is a clean-room restatement, in plain C, of the arithmetic of lukedan/ReSTIR-Vulkan's
 * compute-shader path and of its host-side builders. It exists to check the HIP product path:
 * only tests, the smoke entry point and the CPU-baseline leg of the benchmark may load it, and
 * never as the thing measured or shipped.
 *
 * FP32 functions follow the reference's operation order with the operation semantics fixed as:
 * a/b and sqrt correctly rounded, no FMA contraction, GLSL min/max/clamp as written in the GLSL
 * spec (min(x,y) = y<x?y:x, max(x,y) = x<y?y:x), dot = (x*x'+y*y')+z*z', normalize(v) =
 * v/sqrt(dot(v,v)) per component, pow(x,2.0) = x*x, mix(x,y,a) = x*(1-a)+y*a,
 * sin/cos evaluated in double by a fixed Taylor polynomial and rounded to float (see
 * oracle_sincos), round() = half away from zero.
 *
 * Parity pins (each checked in tests/):
 *   PCG32                 known-answer vector of the PCG reference demo + literal restatement
 *   alias table           Vose invariant sum_j contributions == oriProb, hand-worked FIFO case
 *   BVH build             structural invariants, hand-worked small cases, SAH split brute force
 *   shadow traversal      brute-force FP64 ray/triangle over all triangles (hit/miss away from
 *                         edges), order-independent visit count for misses
 *   BRDF / p_hat          closed forms at special angles, FP64 restatement
 *   reservoir ops         closed-form merges, selection frequency
 *   passes A/B/C          invariants (k=0 identity, static camera, zeroed history), unbiasedness
 *                         of the converged image against the FP64 ground truth
 *   FP64 ground truth     brute-force FP64 direct lighting on tiny scenes
 */
#ifndef RESTIR_ORACLE_H
#define RESTIR_ORACLE_H

#include <stdint.h>
#include "../include/restir_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene {
    const restir_aabb_tree_node *nodes;
    uint32_t n_nodes;
    const restir_triangle *tris;
    uint32_t n_tris;
    const restir_point_light *points; /* may be NULL when n_points == 0 */
    uint32_t n_points;
    const restir_tri_light *tri_lights;
    uint32_t n_tri_lights;
    const restir_alias_table_column *alias;
    uint32_t n_alias;
} oracle_scene;

/* rand.glsl:7-32 */
void oracle_pcg32_stream(uint64_t seed, uint64_t seq, uint32_t n, uint32_t *out);
void oracle_pcg32_float_stream(uint64_t seed, uint64_t seq, uint32_t n, float *out);

/* own sin/cos (float in, float out), used where the GLSL calls sin/cos */
void oracle_sincos(float x, float *s, float *c);

/* common.glsl:7-9, disneyBRDF.glsl:82-90, 73-81, restirUtils.glsl:3-55 */
float oracle_luminance(float r, float g, float b);
float oracle_brdf_luminance(float cosIn, float cosOut, float cosHalf, float cosInHalf,
                            float albedoLum, float roughness, float metallic);
void oracle_brdf_color(float cosIn, float cosOut, float cosHalf, float cosInHalf,
                       const float albedo[3], float roughness, float metallic, float out[3]);
float oracle_phat(const float worldPos[3], const float lightPos[3], const float camPos[3],
                  const float normal[3], const float lightNormal[3], int useLightNormal,
                  float albedoLum, float emissionLum, float roughness, float metallic);
void oracle_phat_full(const float worldPos[3], const float lightPos[3], const float camPos[3],
                      const float normal[3], const float lightNormal[3], int useLightNormal,
                      const float albedo[3], const float emission[3], float roughness,
                      float metallic, float out[3]);
double oracle_phat_f64(const double worldPos[3], const double lightPos[3], const double camPos[3],
                       const double normal[3], double albedoLum, double emissionLum,
                       double roughness, double metallic);

/* reservoir.glsl:28-64, on one reservoir; rng state passed as (state, inc) pairs */
void oracle_add_sample(restir_reservoir *res, const float position[3], const float normal4[4],
                       float emissionLum, int32_t lightIdx, float pHat, float sampleP,
                       uint64_t rng[2]);
void oracle_combine(restir_reservoir *self, const restir_reservoir *other, float pHat,
                    uint64_t rng[2]);

/* misc.cpp:418-497 */
int oracle_build_alias_table(const restir_point_light *points, uint32_t n_point,
                             const restir_tri_light *tris, uint32_t n_tri,
                             restir_alias_table_column *out);
/* Extension (no reference counterpart): one table over n_point point lights followed by n_tri
 * triangle lights, weights luminance and luminance * area, built as createAliasTable builds. A
 * scene whose alias table has n_point + n_tri columns (both > 0) samples from this union. */
int oracle_build_alias_table_mixed(const restir_point_light *points, uint32_t n_point,
                                   const restir_tri_light *tris, uint32_t n_tri,
                                   restir_alias_table_column *out);

/* Seeded k-means light clusters with per-cluster Vose tables (no reference counterpart; the
 * learned proposal's cluster structure). Pinned by k-means invariants in tests/. */
int oracle_build_light_clusters(const restir_point_light *pts, uint32_t n, uint32_t k,
                                uint64_t seed, uint32_t iterations, int32_t *light_cluster,
                                float *light_pwithin, int32_t *cl_begin, int32_t *cl_count,
                                int32_t *cl_lights, restir_alias_table_column *cl_alias,
                                float *centroids);

/* aabbTreeBuilder.cpp:52-214 */
int oracle_build_aabb_tree(const float *tri_xyz, uint32_t n_tris, restir_aabb_tree_node *nodes,
                           restir_triangle *tris, uint32_t *depth_out);

/* softwareRaytracing.glsl:9-85 + visibilityTest.glsl:1-30 (FP32, reference loop semantics) */
int oracle_ray_aabb(const float origin[3], const float dir[3], const float mn[3],
                    const float mx[3]);
int oracle_ray_triangle(const restir_triangle *tri, const float origin[3], const float dir[3]);
void oracle_trace_shadow(const oracle_scene *s, const float *p1, const float *p2, uint32_t n,
                         uint32_t *shadowed, uint32_t *node_visits, uint32_t *tri_tests);
/* FP64 brute force over every triangle of the same shrunk segment; edge_dist[i] receives the
 * smallest barycentric distance to an edge among triangles the segment crosses (or 1). */
void oracle_trace_brute_f64(const restir_triangle *tris, uint32_t n_tris, const float *p1,
                            const float *p2, uint32_t n, uint32_t *shadowed, double *edge_dist);
/* Number of nodes whose box the segment enters, counted recursively (order independent). */
void oracle_count_reachable(const oracle_scene *s, const float *p1, const float *p2, uint32_t n,
                            uint32_t *count);

/* Passes (restirOmni.glsl:86-212, spatialReuse.comp:30-85, unbiasedReuse.glsl:50-185) */
void oracle_pass_a(const oracle_scene *s, const restir_uniforms *u, const restir_gbuffer *cur,
                   const restir_gbuffer *prev, const restir_reservoir *prev_res,
                   restir_reservoir *out);
/* Pass A with RIS over given candidates (the learned proposal's): candidate i of pixel p is
 * (cand_light, cand_q)[i * W*H + p], weighted p_hat / q. */
void oracle_pass_a_candidates(const oracle_scene *s, const restir_uniforms *u, const restir_gbuffer *cur,
                              const restir_gbuffer *prev, const restir_reservoir *prev_res,
                              const int32_t *cand_light, const float *cand_q, restir_reservoir *out);
void oracle_pass_b(const oracle_scene *s, const restir_uniforms *u, int iter,
                   const restir_gbuffer *g, const restir_reservoir *in, restir_reservoir *out);
void oracle_pass_c(const oracle_scene *s, const restir_uniforms *u, int k, int iter,
                   const restir_gbuffer *g, const restir_reservoir *in, restir_reservoir *out);
float oracle_cos_threshold(float degrees);
/* lighting.frag:49-71, 103 */
void oracle_shade(const oracle_scene *s, const float cam[3], float gamma, const restir_gbuffer *g,
                  const restir_reservoir *res, float *rgb_planes, uint64_t rgb_plane_stride);
/* The same with the visibility re-test of the final shading (the 1080p workload's setting): a
 * pixel whose sample has nonzero W and is shadowed from it shades to 0. */
void oracle_shade_visibility(const oracle_scene *s, const float cam[3], float gamma, const restir_gbuffer *g,
                             const restir_reservoir *res, float *rgb_planes, uint64_t rgb_plane_stride);

/* Primary-ray G-buffer by brute force over all triangles (closest t, ties to lower index). */
void oracle_primary_gbuffer(const restir_triangle *tris, uint32_t n_tris,
                            const restir_camera *cam, const int32_t *tri_material,
                            const float *materials, float *position, float *normal,
                            float *albedo, float *material, float *depth, uint64_t plane_stride,
                            uint32_t width, uint32_t height);
void oracle_projection_view(const restir_camera *cam, float aspect, float out_col_major[16]);

/* FP64 ground truth of the direct lighting the image converges to: for every pixel,
 * sum over point lights of f(y) * V(y) in double, V by FP64 brute force (n_tris small) or by the
 * BVH in double (use_bvh). Triangle lights are integrated by n_tri_samples stratified samples.
 * Only pixels with mask[i] != 0 are computed when mask is non-NULL. */
void oracle_ground_truth_f64(const oracle_scene *s, const float cam[3], const restir_gbuffer *g,
                             int use_bvh, uint32_t n_tri_samples, const uint8_t *mask,
                             double *rgb_planes, uint64_t rgb_plane_stride);

#ifdef __cplusplus
}
#endif
#endif
