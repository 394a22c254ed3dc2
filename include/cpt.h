/*
 * cpt.h -- C ABI of the MI355X path tracer (libcpt.so).
 *
 * This is the drop-in boundary for cpp-pt's PathtracerMIS hot path.  The
 * reference has no C ABI of its own: its entry point is the C++ plugin
 * interface `Renderer::render(Scene&, std::string)` (src/renderer/renderer.h:30-33),
 * constructed from `RenderSettings` (renderer.h:11-28) by the factory in
 * src/main.cpp:102-125.  The host C++ class PathtracerMISGPU (cpp-pt_amd/host)
 * implements that interface on top of the functions below; every function
 * here replaces one piece of what `PathtracerMIS::render` and the objects it
 * calls do on the CPU.  Each declaration names the reference code it stands in
 * for.
 *
 * Conventions: plain pointers and sizes only; the caller owns every host
 * buffer and may free it when a call returns; the context owns all device
 * memory.  Every function returns 0 on success or a negative CPT_E* code and
 * never throws or aborts; cpt_last_error() returns the message of the last
 * failure on that context.  A context is used from one thread at a time.
 */
#ifndef CPT_H
#define CPT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPT_OK 0
#define CPT_E_INVALID (-1)   /* bad argument or scene (e.g. out-of-range index) */
#define CPT_E_HIP (-2)       /* HIP runtime error */
#define CPT_E_NOSCENE (-3)   /* call needs cpt_upload_scene first */
#define CPT_E_NODEVICE (-4)  /* no usable gfx950 device */
#define CPT_E_LIMIT (-5)     /* a size limit of the implementation was exceeded */

/* material types (src/bxdf/standard_material.h:46-61, emissive_material.h:28-30) */
#define CPT_MAT_DIFFUSE 0    /* StandardMaterial with metal < 0.1 -> DiffuseBxDF */
#define CPT_MAT_GLOSSY 1     /* StandardMaterial with metal >= 0.1 -> GlossyBxDF(roughness) */
#define CPT_MAT_EMISSIVE 2   /* EmissiveMaterial -> EmissiveBxDF(colour*intensity) */
#define CPT_MAT_DIELECTRIC 3 /* RefractionBxDF(ior) (refraction_bxdf.h), made reachable */
#define CPT_MAT_STANDARD 4   /* StandardMaterial with an image metal texture: DiffuseBxDF or
                                GlossyBxDF chosen per hit (standard_material.h:46-61) */

#define CPT_MATERIAL_FLOATS 16
#define CPT_CAMERA_FLOATS 16

typedef struct cpt_ctx cpt_ctx;

/* Flattened scene.  Replaces the Scene/SceneData object graph
 * (src/scene.h:16-26) that Loader::load_scene builds (src/loader/loader.h:314-374).
 *  - positions/normals/bitangents: n_vertices x 3 floats (Mesh vertex data,
 *    src/shape/mesh.h:101-106; bitangent = cross(normal, tangent), loader.h:237)
 *  - indices: n_triangles x 3 int32 into the vertex arrays; triangle id = its
 *    position in load order (mesh order, then face order: loader.h:353-360)
 *  - tri_material: n_triangles int32 into the material table
 *  - materials: n_materials x 16 floats:
 *      [0] type  [1..3] albedo  [4] roughness  [5..7] emission (colour*intensity)
 *      [8] ior   [9..11] constant normal-map texel  [12] has normal map (1/0)
 *  - light_triangles: the LightGroup in its order (light_group.h:57-60): point
 *    lights as they appear in the .ptscene (loader.h:129-141), then every
 *    emissive triangle (loader.h:362-364).  An entry >= 0 is a triangle id, an
 *    entry e < 0 is point light -1-e.
 *  - camera: o[3], frame columns x[3] y[3] z[3], tan_fovy, aspect, resX, resY
 *    (CameraPerspective, src/camera/camera_perspective.h:19-27)
 *  - env: constant radiance returned for rays that leave the scene
 *    (PathtracerMIS::render_sky returns 0, pathtracerMIS.h:20-22)
 *  - image textures (all optional; n_textures = 0 without them), ImageTexture
 *    (src/texture/image_texture.h:24-40) over Image (src/image/rgb_image.h:161-197):
 *      uvs: n_vertices x 2 floats (the .ptobj texture coordinates, loader.h:236)
 *      textures: n_textures x 4 int32: width, height, channels (1 or 3), offset of
 *        the first texel in `texels`
 *      texels: linear floats (sRGB already decoded where the reference decodes it),
 *        row 0 at the top, channels interleaved
 *      material_textures: n_materials x 4 int32: albedo, roughness, metal and
 *        normal-map textures (-1: the material record's constant)
 *  - point_lights: n_point_lights x 6 floats, position and the radiance that
 *    PointLight::connect_eye_path returns, colour*intensity/(4 pi) computed as the
 *    reference does (point_light.h:20-28); NEE treats its normal as (1,0,0)
 */
typedef struct {
    int32_t n_vertices;
    const float* positions;
    const float* normals;
    const float* bitangents;
    int32_t n_triangles;
    const int32_t* indices;
    const int32_t* tri_material;
    int32_t n_materials;
    const float* materials;
    int32_t n_lights;
    const int32_t* light_triangles;
    float camera[CPT_CAMERA_FLOATS];
    float env[3];
    int32_t flags; /* CPT_SCENE_* */
    const float* uvs;
    int32_t n_textures;
    const int32_t* textures;
    const float* texels;
    const int32_t* material_textures;
    int32_t n_point_lights;
    const float* point_lights;
} cpt_scene_desc;

/* skip computing the reference BVH's test order (used only to break exact-t ties
 * the way the reference does; without it ties go to the lowest triangle id) */
#define CPT_SCENE_NO_TIE_RANK 1
/* build the BVH on the device (PLOC clustering of Morton-sorted triangles,
 * collapsed to the same 4-wide layout) instead of the host's binned SAH build;
 * intersections are identical, only build time and traversal speed differ */
#define CPT_SCENE_GPU_BVH 2

/* One render.  Replaces PathtracerMIS::render (pathtracerMIS.h:135-167):
 * width x height film, spp_axis^2 stratified samples per pixel (the settings
 * file's `spp` is per axis, pathtracerMIS.h:148-149), at most max_depth
 * extension rays per path (the reference's hard cap is 32, pathtracerMIS.h:36).
 * Random numbers: one PCG32 stream per path, seeded from (seed, path id). */
typedef struct {
    int32_t width, height;
    int32_t spp_axis;
    int32_t max_depth;
    uint64_t seed;
    int32_t flags;          /* CPT_RENDER_* */
    int32_t max_paths;      /* paths in flight per pass; 0 = automatic */
    int32_t block_size;     /* threads per block of the wavefront kernels; 0 = automatic */
    int32_t grid_blocks;    /* blocks of the grid-stride kernels; 0 = automatic */
    int32_t view;           /* view id mixed into the seed (multi-view batches) */
    int32_t shade_block;    /* threads per block of the shading kernel (64..1024); 0 = automatic (256) */
    int32_t refill;         /* traversal: idle lanes of a wave (1..64) that make it fetch new rays; 0 = automatic */
    int32_t reserved;
} cpt_render_params;

#define CPT_RENDER_NO_GRAPH 1       /* launch kernels one by one instead of a captured hipGraph */
#define CPT_RENDER_KERNEL_TIMING 2  /* hipEvents around every kernel (implies NO_GRAPH) */
#define CPT_RENDER_COUNT 4          /* count BVH node visits and triangle tests */
#define CPT_RENDER_ACCUMULATE 8     /* add to the film instead of clearing it first */
#define CPT_RENDER_STREAM 16        /* streaming wavefront (path regeneration) instead of passes;
                                       the film is summed with FP64 atomics (order not fixed) */
#define CPT_RENDER_GUIDED 32        /* sample directions from the learned guide mixed with the BSDF
                                       (needs cpt_guide_init) */
#define CPT_RENDER_GUIDE_TRAIN 64   /* with GUIDED: record training samples and train after every pass
                                       (passes default to one sample per pixel) */
/* Integrator (default PathtracerMIS, src/renderer/pathtracerMIS.h:29-130): */
#define CPT_RENDER_NEE_ONLY 128     /* Pathtracer (src/renderer/pathtracer.h:27-110): NEE without MIS,
                                       emitters count only after delta or camera bounces */
#define CPT_RENDER_PURE_PT 256      /* PurePt (src/renderer/pure_pt.h:27-67): BSDF sampling only,
                                       every emitter hit counts, no shadow rays */
#define CPT_RENDER_PPG 512          /* PPG (src/renderer/ppg.h:145-205): the SD-tree guided renderer's
                                       schedule of 1, 2, 4, ... spp iterations, the tree refined between
                                       them, the film holding the last iteration */
#define CPT_RENDER_PPG_STEP 1024    /* one PPG iteration with the current SD-tree (ppg.h:41-143):
                                       spp_axis = samples per pixel (whole-pixel jitter); splats
                                       accumulate until cpt_ppg_update */
#define CPT_RENDER_CORRECTED 4096   /* switch off the reference's biased quirks (every integrator):
                                       MIS weights from the BSDF pdf before Russian roulette
                                       (pathtracerMIS.h:54,123), point lights weighted 1 in NEE
                                       (their (1,0,0) normal, point_light.h:25), survival
                                       probability clamped at 1 (pathtracerMIS.h:24-27), normal
                                       maps at every hit (intersection.h:26-31), and the light
                                       tracer's aspect ratio (camera_perspective.h:75-79) */
#define CPT_RENDER_LIGHT_TRACE 2048 /* Lighttracer (src/renderer/lighttracer.h:17-154): width x height x
                                       spp_axis^2 paths start on the lights, every non-delta vertex
                                       is connected to the camera and splatted into the film (the
                                       reference's image, aspect factor left out as it does);
                                       max_depth bounces (the reference fixes 32); no sums of
                                       squares; one view, no streaming */

typedef struct {
    double seconds;           /* wall time of the render call (device synchronised) */
    int64_t paths;
    int64_t extension_rays;
    int64_t shadow_rays;
    int64_t node_visits;      /* CPT_RENDER_COUNT only */
    int64_t triangle_tests;   /* CPT_RENDER_COUNT only */
    int64_t shadow_node_visits;
    int64_t shadow_triangle_tests;
    double ms_extend, ms_shade, ms_shadow, ms_other;  /* CPT_RENDER_KERNEL_TIMING only */
    int64_t n_extend, n_shade, n_shadow;              /* launches timed */
    int32_t passes;
    int32_t paths_per_pass;
    int32_t block_size;
    int32_t grid_blocks;
    int64_t guide_samples;    /* CPT_RENDER_GUIDE_TRAIN: training samples used */
    int32_t guide_steps;      /* CPT_RENDER_GUIDE_TRAIN: Adam steps taken */
    int32_t graph_captured;   /* 1 if this call captured and instantiated a new pass graph */
    double ms_guide;          /* CPT_RENDER_KERNEL_TIMING, guided renders: guide inference */
    int64_t n_guide;
    double ms_train;          /* CPT_RENDER_KERNEL_TIMING with GUIDE_TRAIN: sample collection and
                                 Adam steps between passes (ABI 1.1) */
    int64_t n_train;          /* passes whose training was timed */
} cpt_render_stats;

/* Create a context on HIP device `device`.  Fails with CPT_E_NODEVICE when the
 * device is not gfx950. */
int cpt_create(cpt_ctx** out, int device);
void cpt_destroy(cpt_ctx* ctx);
const char* cpt_last_error(cpt_ctx* ctx);
/* ABI version (major*1000 + minor) */
int cpt_version(void);
/* Fault injection for tests; no counterpart in the reference.  The context's
 * next `after` device allocations succeed and the one after them fails as if
 * hipMalloc had failed (after < 0: off).  Lets a test check that a call after a
 * failed allocation reallocates instead of using a stale buffer. */
int cpt_debug_fail_alloc(cpt_ctx* ctx, int32_t after);
/* Test hook; no counterpart in the reference.  Converts one 4-wide node as the
 * traversal stores it: node4 = 32 words (low x, high x, low y, high y, low z,
 * high z planes of the 4 children as floats, then 4 child references, 0x7fffffff
 * for an absent child, then 4 unused), qnode_out = 16 words (low corner x y z as
 * floats, the three scale exponents + 127 in bytes 0-2, the 4 child references,
 * then per axis the 4 children's low and high plane bytes, 2 unused).  Host code
 * only: needs no context and no GPU. */
int cpt_debug_quantize_node(const uint32_t* node4, uint32_t* qnode_out);

/* Upload a scene, build the 4-wide BVH and (unless CPT_SCENE_NO_TIE_RANK) the
 * reference test-order table.  Replaces Loader::load_scene's BVH construction
 * (loader.h:353-372 -> BVH::build, src/primitive/bvh.h:151-217) and the
 * LightGroup population (loader.h:362-364). */
int cpt_upload_scene(cpt_ctx* ctx, const cpt_scene_desc* desc);

/* Host only (no device needed): parse a scene file of the reference's format
 * and flatten it into *desc.  Replaces Loader::load_scene's parsing
 * (src/loader/loader.h:314-351: camera, obj + mtl and light lines, read by
 * parse_camera / parse_light / parse_mat / parse_obj, loader.h:117-254) for a caller that has no
 * flattened scene of its own: the reference's renderer receives only `Scene&`
 * (renderer.h:30-33), whose camera and materials keep their data private, so
 * its adapter loads `RenderSettings::scene_name` again here.  The camera is set
 * up for a res_x x res_y film.  *file owns every array *desc points to; release
 * it with cpt_scene_file_free after cpt_upload_scene (which copies).  On failure
 * returns CPT_E_INVALID, sets *file to NULL and writes the reason into err
 * (err_len bytes including the terminating zero; err may be NULL). */
typedef struct cpt_scene_file cpt_scene_file;
int cpt_scene_file_load(const char* path, int32_t res_x, int32_t res_y, cpt_scene_file** file,
                        cpt_scene_desc* desc, char* err, int32_t err_len);
void cpt_scene_file_free(cpt_scene_file* file);

/* BVH statistics of the uploaded scene (6 entries): [nodes, leaves, max depth,
 * triangles, build time in microseconds, 1 if built on the device else 0] */
int cpt_bvh_info(cpt_ctx* ctx, int64_t* info6);
/* The reference BVH's test order of every triangle (n_triangles int32). */
int cpt_tie_rank(cpt_ctx* ctx, int32_t* rank_out);

/* Host only (no device needed): the reference BVH's test order for a triangle
 * soup, as cpt_upload_scene computes it (bvh.h:57-217 restated). */
int cpt_reference_tie_rank(const float* positions, const int32_t* indices, int32_t n_triangles, int32_t* rank_out);

/* Closest hit for n rays (n x 7 floats: ox oy oz dx dy dz tmax, host memory).
 * Replaces BVH::intersect (bvh.h:219-221 -> BVHNode::intersect 22-37 ->
 * Triangle::intersect, src/shape/triangle.h:18-116).  tri_out = triangle id or
 * -1, t_out = hit distance (bit-identical to the reference's t). */
int cpt_trace_closest(cpt_ctx* ctx, const float* rays, int32_t n, int32_t* tri_out, float* t_out);
/* Any hit (occlusion) for n rays.  Replaces BVH::intersect_any (bvh.h:222-224,
 * triangle.h:117-179). occluded_out: 1 if any triangle has t in [EPS, tmax]. */
int cpt_trace_any(cpt_ctx* ctx, const float* rays, int32_t n, uint8_t* occluded_out);
/* Same on device-resident buffers (rays: n x 8 floats ox oy oz tmax dx dy dz 0;
 * hits: n x 4 words t, beta, gamma, triangle id) on stream `stream` (NULL = the
 * context's stream); no host synchronisation. */
int cpt_trace_closest_device(cpt_ctx* ctx, const void* d_rays, int32_t n, void* d_hits, void* stream);
/* cpt_trace_closest_device without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernel, same results bit for bit; its traversal overflow stacks live in a
 * file-scope device array of 52 MiB per GPU (1024 blocks x 128 threads x 104
 * entries, located with hipGetSymbolAddress when a context is created), so the
 * launch is capped at 1024 blocks, each wave walking its share of the batch; any
 * n >= 0 is accepted.  Every call on one GPU uses that GPU's array: calls that
 * overlap on one GPU (other streams, other contexts) race and must be serialised
 * by the caller.  CPT_E_LIMIT if the scene's BVH needs deeper stacks than the
 * array holds (scene upload already refuses such trees). */
int cpt_trace_closest_device_noalloc(cpt_ctx* ctx, const void* d_rays, int32_t n, void* d_hits, void* stream);

/* Render into the context's device film (FP64 accumulators).  Replaces the
 * pixel loop of PathtracerMIS::render (pathtracerMIS.h:139-160). */
int cpt_render(cpt_ctx* ctx, const cpt_render_params* params, cpt_render_stats* stats);

/* SD-tree of the PPG arm (src/algorithms/sdtree.h).
 * cpt_ppg_reset: a fresh tree over the scene's triangle bounds (SDTree ctor, sdtree.h:434-438).
 * cpt_ppg_update: add the splats of the PPG steps since the last update, then refine
 *   (SDTree::update / subdivide, sdtree.h:454-504; DTree::adapt_and_build_leaves 172-233).
 * cpt_ppg_export: the flattened tree the next step uses and the current splat sums:
 *   sizes = {spatial nodes, directional trees, quadtree nodes, sampling leaves, accumulators};
 *   snode (>= 0: children v, v+1; < 0: dtree -1-v), dinfo (4 per dtree: quad base, leaf base,
 *   leaf count, 0), quad (>= 0: first child; < 0: accumulator -1-v), rect (4 floats per leaf:
 *   phi, cos theta, sizes), sums/counts per accumulator, bounds = lo[3], hi[3].  Any array
 *   pointer may be NULL (sizes first). */
int cpt_ppg_reset(cpt_ctx* ctx);
int cpt_ppg_update(cpt_ctx* ctx);
int cpt_ppg_export(cpt_ctx* ctx, int32_t* sizes, int32_t* snode, int32_t* dinfo, int32_t* quad, float* rect,
                   float* sums, uint32_t* counts, float* bounds);

/* Launch configuration found by cpt_autotune. */
typedef struct {
    int32_t block_size, grid_blocks, shade_block, refill;  /* grid_blocks 0 = resident grid */
    int32_t trials;          /* configurations timed */
    int32_t reserved0;
    double samples_per_s;    /* best trial */
    double seconds;          /* wall time of the search */
    int32_t reserved[4];
} cpt_tune_result;

/* Time short renders of the uploaded scene at params' size, depth and flags over
 * candidate launch configurations (block size, traversal grid, refill threshold,
 * shading block; coordinate descent from the defaults) and keep the fastest: later
 * cpt_render calls on the same scene, size, depth and guided flag use it for every
 * launch parameter they leave 0.  The film holds the last trial afterwards.
 * No counterpart in the reference (its renderer has no launch parameters). */
int cpt_autotune(cpt_ctx* ctx, const cpt_render_params* params, cpt_tune_result* out);
/* Copy the film out: rgb (w*h*3 floats) = sum / spp_axis^2 of the last
 * render, row j = 0 at the top like RgbImage::put_pixel (rgb_image.h:208-210);
 * sums (w*h*3 doubles, may be NULL) = the raw FP64 accumulators; sumsq (may be
 * NULL) = per-pixel sums of squared path radiance (variance estimates). */
int cpt_read_film(cpt_ctx* ctx, float* rgb, double* sums, double* sumsq);

/* Multi-view batch in one launch sequence: n_views cameras (16 floats each, as
 * cpt_scene_desc.camera) with their own width/height/spp_axis.  Paths of all
 * views share the wavefront queues.  films: concatenated w*h*3 float images in
 * view order.  view_ms (may be NULL): per-view completion time in ms from the
 * start of the batch. */
int cpt_render_views(cpt_ctx* ctx, int32_t n_views, const float* cameras, const int32_t* wh_spp,
                     int32_t max_depth, uint64_t seed, int32_t flags, float* films, double* view_ms,
                     cpt_render_stats* stats);

/* ---- learned path guiding -------------------------------------------------
 * Replaces the SD-tree of the reference's practical-path-guiding renderer
 * (src/renderer/ppg.h:41-214, src/algorithms/sdtree.h:445-538): a directional
 * distribution per path vertex, here an MLP (one-blob encoding of position,
 * normal and outgoing direction, 7 x 32 bins -> 4 x 64 ReLU -> 8 von Mises-Fisher
 * lobes, or 4 x 128 -> 16 lobes) with BF16 weights on MFMA and FP32 accumulation.
 * SDTree::sample/pdf -> the mixture's sample/pdf inside the guided shade kernel;
 * SDTree::splat/update -> training samples and cpt_guide_train_step (forward,
 * backward, Adam on the GPU).
 * Parameters (width H, O outputs = 48 for 8 lobes, 80 for 16): layer l weights W_l
 * (K_l x N_l row-major, K = 224, H, H, H, H; N = H, H, H, H, O) each followed by
 * its bias (N_l): 30000 floats at width 64, 88656 at width 128.  Outputs per
 * vector: 5 per lobe (logit, log kappa, unnormalised mean xyz), padded to O. */
typedef struct {
    float lr;              /* Adam step size (0 = 1e-3) */
    float beta1, beta2;    /* 0 = 0.9, 0.99 */
    float eps;             /* 0 = 1e-8 */
    uint64_t seed;         /* weight initialisation */
    float guide_fraction;  /* probability of sampling the guide (0 = 0.5) */
    int32_t train_paths;   /* guided renders: paths per pass that record training samples (0 = 65536) */
    int32_t train_steps;   /* guided renders: Adam steps after each pass (0 = 1) */
    int32_t guide_depth;   /* guided renders: guide the vertices of the first this many bounces (0 = all) */
    int32_t guide_target;  /* training target: 0 = the product L_i f cos (default), 1 = the incident
                              radiance L_i alone (what PPG splats, ppg.h:130-141) */
    int32_t infer_fp8;     /* 1: guided renders and cpt_guide_forward run the network in FP8 (e4m3
                              weights with per-layer scales, activation scales per 16 samples, FP8 MFMA,
                              FP32 accumulation); training stays BF16 */
    int32_t width;         /* hidden width: 64 (8 lobes; 0 = 64) or 128 (16 lobes, the large-scene
                              network: BF16 inference reads its weights through the L2, FP8
                              inference keeps them in LDS) */
    int32_t fraction_ramp; /* guided renders: the guide fraction rises linearly from 0 to
                              guide_fraction over the network's first this many Adam steps, so
                              that an untrained guide is sampled rarely (0 = constant) */
} cpt_guide_params;

int cpt_guide_num_params(void);  /* of the width-64 network */
/* shape4 = [hidden width, lobes, outputs per vector, parameters] of the context's
 * guide (of the width-64 network before cpt_guide_init). */
int cpt_guide_shape(cpt_ctx* ctx, int32_t* shape4);
/* Allocate (first call) and initialise the guide: Glorot-uniform weights from `seed`. */
int cpt_guide_init(cpt_ctx* ctx, const cpt_guide_params* params);
int cpt_guide_get_params(cpt_ctx* ctx, float* params_out);
int cpt_guide_set_params(cpt_ctx* ctx, const float* params_in);
/* Raw network outputs (n x 48, or n x 80 at width 128) for n input vectors (n x 7
 * floats in [0,1]). */
int cpt_guide_forward(cpt_ctx* ctx, const float* inputs, int32_t n, float* outputs);
/* Stored mixtures for n input vectors, computed by the code guided renders run per
 * bounce once they have gathered a hit's inputs (its encode, forward pass in the
 * configured precision, and decode).  mixtures_out: 3 x lobes x n 32-bit words,
 * value v of sample s at v * n + s, in blocks of `lobes`: for lobe k, half2
 * (kappa mu_x, kappa mu_y) and half2 (kappa mu_z, w~) as raw bits (half precision,
 * low half first; kappa = exp(clamped log kappa), mu the unit mean, w~ the softmax
 * weight), then the float c_k = log(w~_k / S C(kappa'_k)) - kappa'_k with S the sum of
 * the w~, kappa' = |kappa mu| from the halves and C(k) = k / (2 pi (1 - exp(-2 k))).
 * The mixture pdf of a direction w is sum_k exp(kappa mu_k . w + c_k). */
int cpt_guide_mixtures(cpt_ctx* ctx, const float* inputs, int32_t n, float* mixtures_out);
/* Test hook, no counterpart in the reference: the stored mixtures of queue entries
 * 0 .. n-1 of the last guided bounce of the last guided render (after a render of
 * max_depth 1 and one sample per pixel, entry e is pixel e's camera-ray hit), laid
 * out as cpt_guide_mixtures' output.  Lets a test check the per-bounce gather of the
 * hit's inputs as well as the network. */
int cpt_debug_guide_mixtures(cpt_ctx* ctx, int32_t n, float* mixtures_out);
/* One training step on n samples (n x 11 floats: 7 inputs, direction xyz, weight =
 * target radiance / sampling pdf).  Loss = sum_i w_i (-log q(dir_i)) / sum_i w_i.
 * loss_out (may be NULL) = that loss; grad_out (may be NULL, one float per
 * parameter) = its gradient; apply != 0 also takes the Adam step. */
int cpt_guide_train_step(cpt_ctx* ctx, const float* samples, int32_t n, int32_t apply, float* loss_out,
                         float* grad_out);

#ifdef __cplusplus
}
#endif
#endif
