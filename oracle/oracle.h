/*
 * CPU oracle for the cpp-pt PathtracerMIS hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (libcpt.so and everything under cpp-pt_amd/) never links, loads or
 * calls it.
 *
 * It restates, in plain C++ with the reference's own float/double promotion
 * rules, the algorithms of (paths relative to the cpp-pt source tree):
 *   src/primitive/bvh.h:57-229        binary BVH build (5 centroid bins, area cost) + traversal
 *   src/primitive/aabb.h:19-61        slab test (max_t effectively ignored)
 *   src/primitive/simple_group.h:22-39, primitive_leaf.h:27-37
 *   src/shape/triangle.h:18-269       Cramer ray/triangle test, area, uniform sampling
 *   src/shape/intersection.h:25-33    first-hit normal mapping
 *   src/camera/camera_perspective.h:19-39
 *   src/bxdf/{bxdf,diffuse_bxdf,emissive_bxdf,glossy_bxdf,refraction_bxdf}.h
 *   src/math/microfacet_distribution.h:13-165, src/math/math.h:647-750
 *   src/light/{light_group,shape_light}.h
 *   src/renderer/pathtracerMIS.h:20-167
 *
 * Parity pins (tests/test_oracle_*.py): brute-force closest hit, closed-form
 * Lambertian and GGX normalisation, triangle area-uniform density, the
 * all-emissive furnace (image == Le exactly), Lambert's polygon formula for
 * one-bounce irradiance, mt19937 known answer. Functions without such a pin are
 * marked "parity unpinned" beside their definition.
 */
#ifndef CPT_ORACLE_H
#define CPT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* material record: 16 floats
 * [0] type (0 diffuse, 1 glossy GGX, 2 emissive, 3 dielectric)
 * [1..3] albedo rgb   [4] roughness   [5..7] emission rgb (colour*intensity)
 * [8] ior   [9..11] constant normal-map texel (rgb, 0.5,0.5,1 = flat)
 * [12] has_normal_map (1 for StandardMaterial, 0 for emissive)  [13..15] unused */
void* orc_scene_create(int nv, const float* pos, const float* nrm, const float* bitan,
                       int nt, const int32_t* idx, const int32_t* tri_mat,
                       int nm, const float* mats,
                       int nl, const int32_t* light_tris,
                       const float* cam /* 16: o[3], x[3], y[3], z[3], tan_fovy, aspect, resX, resY */,
                       const float* env /* 3 */);
/* image textures (loader.h:45-97): per-vertex uv (nv x 2), images (n x 4: width,
 * height, channels 1|3, first texel), linear texels, per material the image of
 * its albedo, roughness, metal and normal-map slot (nm x 4, -1 = the record's
 * constant).  A metal image makes the Lambertian/GGX choice per hit. */
void orc_scene_set_textures(void* s, const float* uvs, int n_images, const int32_t* images, int n_texels,
                            const float* texels, const int32_t* material_textures);
/* 1: switch off the reference's biased quirks (the library's CPT_RENDER_CORRECTED) */
void orc_set_corrected(void* s, int corrected);
/* Lighttracer (lighttracer.h:17-154): W*H*S*S light paths splatted into sum_out
 * (H x W x 3, spp times the image); sampler 0 per-path PCG, 1 per-column mt19937 */
int orc_render_lt(void* s, int W, int H, int S, int max_depth, int sampler, uint64_t seed, double* sum_out,
                  int nthreads, int64_t* counts);
/* point lights (n x 6: position, radiance); light_tris entries -1-k refer to them */
void orc_scene_set_point_lights(void* s, int n, const float* rec);
/* per ray: the closest hit's material after its textures (10 floats: type or -1
 * for a miss, albedo rgb, roughness, normal-map texel rgb, u, v) */
void orc_hit_material(void* s, const float* rays, int n, float* out);
/* Image<T>::get_pixel(i, j) with the REPEAT border (channels 1 or 3 floats) */
void orc_image_texel(void* s, int img, int i, int j, float* out);
void orc_scene_destroy(void* s);

/* reference BVH (bvh.h:151-217) statistics and DFS test order of every triangle */
int orc_bvh_nodes(void* s);
void orc_bvh_rank(void* s, int32_t* rank_out);

/* rays: n x 7 floats (ox oy oz dx dy dz tmax). mode 0 = reference BVH
 * traversal, 1 = brute force over all triangles with the reference's tie rule. */
void orc_trace_closest(void* s, const float* rays, int n, int mode, int32_t* tri, float* t,
                       int nthreads);
void orc_trace_any(void* s, const float* rays, int n, int mode, uint8_t* occluded, int nthreads);

/* camera ray for film coordinates in [-1,1]^2 (camera_perspective.h:30-39) */
void orc_camera_ray(void* s, float cx, float cy, float* o3, float* d3);

/* Path tracing.  sampler 0 = per-path PCG32 (the GPU's stream), 1 = per-column
 * std::mt19937 + uniform_real_distribution<float> (the reference's RandomSampler;
 * column i uses seed i + (seed & 0x7fff)*W, so seed 0 is the reference exactly).
 * Renders columns [col0, col1) and rows [row0, row1) of a W x H film with spp_axis^2
 * stratified samples per pixel.  sum_out (W*H*3 doubles, may be NULL) receives the
 * per-pixel radiance sums, sumsq_out the sums of squares, film_out (W*H*3 floats, may be NULL) the reference's
 * float average acc/spp^2.  counts (may be NULL) receives
 * [paths, extension rays, shadow rays].  Returns 0. */
int orc_render(void* s, int W, int H, int spp_axis, int max_depth, int sampler, uint64_t seed,
               int col0, int col1, int row0, int row1, double* sum_out, double* sumsq_out,
               float* film_out, int nthreads, int64_t* counts);

/* Integrator of orc_render / orc_path: 0 PathtracerMIS (pathtracerMIS.h:29-130),
 * 1 Pathtracer, NEE without MIS (pathtracer.h:27-110), 2 PurePt (pure_pt.h:27-67). */
void orc_set_integrator(void* s, int integrator);
/* view id of orc_render's per-path PCG ids, (view << 40) + pixel * spp + sample, as
 * the library numbers the paths of view `view` in a multi-view batch (default 0) */
void orc_set_view(void* s, uint32_t view);

/* One iteration of PPG (src/renderer/ppg.h:41-143) with a fixed, flattened SD-tree (the
 * layout of cpt_ppg_export; bounds = lo[3], hi[3]): spp whole-pixel samples per pixel, PCG
 * path id (view << 40) + pixel*spp + sample.  sum_out receives the per-pixel radiance sums,
 * splat_sum / splat_cnt (n_slots each) what the paths splat into each quadtree leaf. */
int orc_ppg_render(void* s, int W, int H, int spp, int max_depth, uint64_t seed, uint32_t view, const int32_t* snode,
                   const int32_t* dinfo, const int32_t* quad, const float* rect, const float* bounds, int n_slots,
                   double* sum_out, double* splat_sum, int64_t* splat_cnt, int nthreads);

/* One path with an explicit PCG path id; returns radiance in rgb3. */
void orc_path(void* s, int W, int H, int spp_axis, int max_depth, uint64_t seed, int pixel,
              int sample, float* rgb3);

/* --- unit-level entry points (fixed random numbers) ---------------------- */
/* Lambertian sample (diffuse_bxdf.h:34-47): n, wo -> wi, pdf */
void orc_diffuse_sample(const float* n, const float* wo, float u1, float u2, float* wi, float* pdf);
/* GGX glossy (glossy_bxdf.h:29-72): frame t,b,n (columns), roughness */
void orc_glossy_sample(const float* t, const float* b, const float* n, float rough, const float* wo,
                       float u1, float u2, float* wi, float* pdf, int* delta);
float orc_glossy_eval(const float* t, const float* b, const float* n, float rough, const float* wo,
                      const float* wi);
float orc_glossy_pdf(const float* t, const float* b, const float* n, float rough, const float* wo,
                     const float* wi);
/* microfacet D(wh) and pdf(wo,wh) in the local frame (microfacet_distribution.h:121-165) */
float orc_ggx_d(float alpha, const float* wh);
float orc_ggx_pdf(float alpha, const float* wo, const float* wh);
void orc_ggx_sample(float alpha, const float* wo, float u1, float u2, float* wh);
/* Uniform triangle sample (triangle.h:209-269): returns position */
void orc_triangle_sample(void* s, int tri, float u1, float u2, float* pos, float* nrm, float* area);

/* PCG32 stream used by both the oracle and the GPU (first n floats of a path) */
void orc_pcg_floats(uint64_t seed, uint64_t path_id, int n, float* out);
/* std::mt19937(seed) + uniform_real_distribution<float>(0,1): first n floats */
void orc_mt_floats(uint32_t seed, int n, float* out);
/* GGX Lambda and G1 in the local frame (microfacet_distribution.h:137-150) */
float orc_ggx_lambda(float alpha, const float* w);
float orc_ggx_g1(float alpha, const float* w);
/* RefractionBxDF::sample (refraction_bxdf.h:26-52): shading normal n, outgoing wo -> wi */
void orc_dielectric_sample(const float* n, const float* wo, float ior, float* wi);
/* SD-tree lookups on the flattened layout of cpt_ppg_export (sdtree.h:268-396 get_leaf,
 * 400-423 sample_direction, 525-538 pdf): pts, dirs n x 3; u3 n x 3 uniforms per sample */
void orc_sd_lookup(const int32_t* snode, const int32_t* dinfo, const int32_t* quad, const float* rect,
                   const float* bounds, int n, const float* pts, int32_t* dtree_out);
void orc_sd_pdf(const int32_t* snode, const int32_t* dinfo, const int32_t* quad, const float* rect,
                const float* bounds, int n, const float* pts, const float* dirs, float* pdf_out);
void orc_sd_sample(const int32_t* snode, const int32_t* dinfo, const int32_t* quad, const float* rect,
                   const float* bounds, int n, const float* pts, const float* u3, float* dirs_out, float* pdf_out);
/* GGX visible-normal samples for n (u1, u2) pairs (microfacet_distribution.h:96-118) */
void orc_ggx_sample_batch(float alpha, const float* wo, int n, const float* u2, float* wh);
/* survival probability of Russian roulette (pathtracerMIS.h:24-27); corrected = clamped at 1 */
float orc_russian_roulette(const float* c, int corrected);

#ifdef __cplusplus
}
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* Guide MLP (FP64; bf16 != 0 rounds to BF16 where the GPU does). */
int orc_mlp_num_params(void);
void orc_mlp_set_width(int width);  /* 64 (8 lobes, default) or 128 (16 lobes) */
void orc_mlp_forward(const float* params, const float* inputs, int n, int bf16, double* outputs);
double orc_mlp_mixture_pdf(const double* outputs48, const double* dir3);
void orc_mlp_mixture_pdf_batch(const double* outputs48, const double* dirs, int n, double* out);
double orc_mlp_loss_grad(const float* params, const float* samples, int n, int bf16, double* grad);
void orc_mlp_adam(float* params, double* m, double* v, const double* grad, int step, double lr, double b1, double b2,
                  double eps);
#ifdef __cplusplus
}
#endif
