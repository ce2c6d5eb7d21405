/* xenon_hip.h — C ABI of the MI355X (gfx950) wavefront path-tracing backend.
 *
 * libxenon_hip.so is the drop-in for the reference's GPU backend contract,
 * CudaRenderer (src/cuda/cuda_renderer.cuh:14-55):
 *
 *   CudaRenderer(w, h, max_bounces, min_bounces)  -> xnh_create
 *   CudaRenderer::upload_scene(const Scene&)      -> xnh_upload_scene
 *   CudaRenderer::render_frame(scene, cam, chain) -> xnh_render (spp_count 1)
 *   CudaRenderer::spp()                           -> xnh_spp
 *   ~CudaRenderer()                               -> xnh_destroy
 *   CUDA_CHECK abort (cuda/cuda_utils.cuh:12-20)  -> negative return codes
 *                                                    + xnh_last_error
 *
 * Numerics follow the reference's CPU oracle, render/cpu_tiled.cpp:13-237
 * (same per-pixel seeds, RNG draw order, BSDF, NEE/MIS and Russian
 * roulette), not the reference's CUDA kernels.
 *
 * Conventions: every int-returning call returns XNH_OK (0) or a negative
 * XNH_E_* code and never aborts; the library copies what it needs and keeps
 * no caller pointer; one context is used by one thread at a time. Images
 * are W*H*3 float32 in pixel order (index y*W+x, y = 0 the bottom row, as
 * the reference's accumulation buffer); the reference's swapchain writer
 * flips rows (cpu_tiled.cpp:228-229) and a wrapper does the same.
 */
#ifndef XENON_HIP_H
#define XENON_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define XNH_OK 0
#define XNH_E_INVALID (-1)     /* bad argument */
#define XNH_E_HIP (-2)         /* HIP runtime error (message in xnh_last_error) */
#define XNH_E_NO_SCENE (-3)    /* render/trace before xnh_upload_scene */
#define XNH_E_OVERFLOW (-4)    /* a traversal stack exceeded its limit */
#define XNH_E_NO_DEVICE (-5)   /* no usable gfx950 device */
#define XNH_E_ALLOC (-6)       /* device or host allocation failed */

typedef struct xnh_ctx xnh_ctx;
typedef struct xnh_scene xnh_scene;          /* host scene builder */
typedef struct xnh_gpu_scene xnh_gpu_scene;  /* a scene resident on the device */

#define XNH_FLAG_NO_GRAPH 1                  /* launch kernels directly, no hipGraph replay */

/* Renderer configuration. width/height/bounces are used as given; the
 * zero-valued tuning fields take defaults. */
typedef struct {
  int32_t width, height;
  int32_t max_bounces;       /* reference default 8 (scene/scene_file.h:16) */
  int32_t min_bounces;       /* Russian roulette from this depth, default 3 */
  int32_t spp_per_launch;    /* samples per pixel in flight per wavefront batch (0: auto) */
  int32_t block_size;        /* work-group size of the traversal kernels (0: auto) */
  int32_t lds_stack;         /* traversal stack entries held in LDS per lane (0: auto) */
  int32_t flags;             /* XNH_FLAG_* */
} xnh_config;

/* Pinhole camera as the reference's Camera exposes it for GPU upload
 * (camera/camera.h:40-43). */
typedef struct {
  float eye[3], lower_left[3], horizontal[3], vertical[3];
} xnh_camera;

/* Binary-identical to the reference's BLASNode and TLASNode
 * (geometry/blas.h:21-28, geometry/tlas.h:91-96). */
typedef struct {
  float lo[3], hi[3];
  uint32_t child;            /* interior: left child (right = left+1); leaf: first slot */
  uint32_t count;            /* 0: interior; >0: leaf of `count` slots */
} xnh_bvh_node;

/* One triangle mesh with its BLAS, as TriangleMesh (geometry/mesh.h:21-31)
 * and BLAS::nodes()/prims() (geometry/blas.h:59-60) hold them. */
typedef struct {
  int32_t num_vertices, num_triangles;
  const float *vx, *vy, *vz;           /* positions, SoA */
  const float *nx, *ny, *nz;           /* vertex normals, SoA (may be NULL: flat) */
  const int32_t* indices;              /* 3 per triangle */
  const int32_t* mat_ids;              /* 1 per triangle */
  int32_t num_nodes;
  const xnh_bvh_node* nodes;
  const uint32_t* prims;               /* leaf slot -> triangle index */
} xnh_mesh;

/* Instance (geometry/tlas.h:76-87): column-major 3x4 matrices as in
 * AffineTransform (geometry/transform.h:24-62). mesh XNH_PRIM_GEOMETRY: the
 * scene's analytic primitives (its PrimBLAS, GeomHandle::from_prim). */
#define XNH_PRIM_GEOMETRY (-1)
typedef struct {
  int32_t mesh;
  int32_t instance_id;
  float local_to_world[12];
  float world_to_local[12];
} xnh_instance;

#define XNH_MAT_CONDUCTOR 1u
#define XNH_MAT_TRANSMISSIVE 2u
#define XNH_MAT_SUBSURFACE 4u
#define XNH_MAT_DELTA 8u
#define XNH_MAT_EMISSIVE 16u

/* Prepared material (material/material.h:19-49 after material_prepare). */
typedef struct {
  float base_color[3];
  float metallic, ior, alpha, alpha_x, alpha_y;
  float f0[3];
  float subsurface;
  float subsurface_color[3];
  uint32_t flags;
} xnh_material;

/* Area light (scene/scene.h:19-24): triangle tri_idx of mesh mesh_id. */
typedef struct {
  uint32_t mesh_id, tri_idx;
  float emission[3];
  float area;
} xnh_light;

/* Directional light (scene/scene.h:43-49); emission = color * intensity,
 * direction toward the sun, unit length. */
typedef struct {
  float direction[3];
  float emission[3];
} xnh_dir_light;

/* Gradient sky (scene/environment.h:26-51); intensity 0 disables it. */
typedef struct {
  float zenith[3], horizon[3], ground[3];
  float sharpness, intensity;
} xnh_sky;

/* Analytic primitive (geometry/primitives.h:54-318), p by type:
 *   XNH_PRIM_SPHERE  center xyz, radius
 *   XNH_PRIM_PLANE   unit normal xyz, offset (dot(normal, point))
 *   XNH_PRIM_BOX     center xyz, half-extents xyz
 *   XNH_PRIM_DISK    center xyz, unit normal xyz, radius
 *   XNH_PRIM_QUAD    corner xyz, edge u xyz, edge v xyz */
#define XNH_PRIM_SPHERE 0
#define XNH_PRIM_PLANE 1
#define XNH_PRIM_BOX 2
#define XNH_PRIM_DISK 3
#define XNH_PRIM_QUAD 4
typedef struct {
  int32_t type, mat_id;
  float p[10];
} xnh_prim;

/* Everything upload_scene reads from the reference's Scene
 * (scene/scene.h:52-75) after build_acceleration. */
typedef struct {
  int32_t num_meshes;
  const xnh_mesh* meshes;
  int32_t num_instances;
  const xnh_instance* instances;
  int32_t num_tlas_nodes;
  const xnh_bvh_node* tlas_nodes;
  const uint32_t* tlas_inst_indices;   /* TLAS::inst_indices() */
  int32_t num_materials;
  const xnh_material* materials;
  int32_t num_lights;
  const xnh_light* lights;
  int32_t num_dir_lights;
  const xnh_dir_light* dir_lights;
  xnh_sky sky;
  /* Scene::all_prims and its PrimBLAS (geometry/prim_blas.h:54-110):
   * nodes as xnh_bvh_node, prim_order = PrimBLAS::prims_ (leaf slot ->
   * primitive index). Zero / NULL: no primitives. */
  int32_t num_prims;
  const xnh_prim* prims;
  int32_t num_prim_nodes;
  const xnh_bvh_node* prim_nodes;
  const uint32_t* prim_order;
} xnh_flat_scene;

typedef struct {
  float org[3], dir[3];
  float tmin, tmax;
} xnh_ray;

/* Closest hit as the reference's HitRecord after TLAS fix-up
 * (math/ray.h:47-59, geometry/tlas.cpp:156-168). Miss: t = +inf, ids -1. */
typedef struct {
  float t, u, v;
  int32_t instance_id, prim_id, mat_id;
  float pos[3], normal[3], geo_normal[3];
  int32_t front_face;
} xnh_hit;

/* Counters of the last xnh_render / xnh_trace_* call. */
typedef struct {
  uint64_t paths;            /* camera paths started */
  uint64_t closest_rays;     /* closest-hit rays traced */
  uint64_t shadow_rays;      /* shadow rays traced */
  uint64_t node_tests;       /* closest-hit BVH boxes tested (BLAS pair = 2, TLAS pop = 1) */
  uint64_t interior;         /* closest-hit interior nodes expanded */
  uint64_t tri_tests;        /* closest-hit ray-triangle tests */
  uint64_t inst_visits;      /* closest-hit instance ray transforms */
  uint64_t shadow_node_tests;   /* the same four counts for shadow (any-hit) rays */
  uint64_t shadow_interior;
  uint64_t shadow_tri_tests;
  uint64_t shadow_inst_visits;
  uint64_t nonfinite;        /* non-finite path radiances (kept, as the reference does) */
  double render_ms;          /* device time of the call (HIP events) */
  double closest_ms;         /* device time inside the closest-hit kernels */
  double shadow_ms;          /* device time inside the shadow kernels */
  double shade_ms;           /* device time inside the shading kernels */
  int32_t closest_launches;  /* closest-hit kernel launches in the call */
  int32_t stack_overflows;   /* lanes whose traversal stack exceeded its limit */
  /* SIMD use of the traversal, with the work counters (XNH_PROFILE_WORK and
   * the ray-batch calls; zero otherwise): wave iterations of the interior
   * phase, of the leaf phase and of the round loop, lane-level leaf steps
   * and rounds. A phase's active-lane fraction is its lane steps (interior,
   * leaf_steps) over 64 x its wave iterations. Closest hit, then any hit. */
  uint64_t interior_iters, leaf_iters, leaf_steps, round_iters, round_lanes;
  uint64_t shadow_interior_iters, shadow_leaf_iters, shadow_leaf_steps, shadow_round_iters, shadow_round_lanes;
} xnh_stats;

/* ---- renderer (CudaRenderer contract) ---------------------------------- */
int xnh_create(const xnh_config* cfg, xnh_ctx** out);
void xnh_destroy(xnh_ctx* ctx);
const char* xnh_last_error(const xnh_ctx* ctx);
int xnh_upload_scene(xnh_ctx* ctx, const xnh_flat_scene* scene);
/* Adds samples spp_first .. spp_first+spp_count-1 (1-based sample indices,
 * the value the reference's spp_ holds while rendering that frame) to the
 * device accumulator. rgb_out (optional, host, W*H*3) receives the running
 * average over all samples added since the last reset. */
int xnh_render(xnh_ctx* ctx, const xnh_camera* cam, int32_t spp_first, int32_t spp_count, float* rgb_out);
/* One render job of a batched call: `scene` (NULL: the scene given to
 * xnh_upload_scene) seen through `camera` at width x height, samples
 * spp_first .. spp_first+spp_count-1, its own depth limits. Outputs are
 * optional host buffers of W*H*3 floats in pixel order: the average and/or
 * the sums of this call's samples. */
typedef struct {
  const xnh_gpu_scene* scene;
  xnh_camera camera;
  int32_t width, height;
  int32_t spp_first, spp_count;
  int32_t max_bounces, min_bounces;
  float* rgb_out;
  float* sum_out;
} xnh_job;

/* Additional device-resident scenes for xnh_render_jobs. No counterpart in
 * the reference (one scene per renderer there). */
int xnh_scene_upload(xnh_ctx* ctx, const xnh_flat_scene* scene, xnh_gpu_scene** out);
void xnh_scene_release(xnh_ctx* ctx, xnh_gpu_scene* scene);
/* Renders many jobs of uneven resolution, sample count and depth together:
 * their samples share the wavefront queues (per-job queue segments) and
 * every batch is one replay of the same captured kernel sequence. Each
 * job's result equals rendering it alone, bit for bit. No counterpart in
 * the reference. */
int xnh_render_jobs(xnh_ctx* ctx, const xnh_job* jobs, int32_t njobs);

/* ---- path guiding (no counterpart in the reference) ---------------------
 * A multiresolution hash grid (16 levels x 2^t_log2 entries x 2 FP16
 * features) and a bias-free 3x64 ReLU MLP map (position, normal) to an
 * 8-lobe von Mises-Fisher mixture over incident direction; FP16 MFMA
 * forward/backward with FP32 accumulation, Adam on FP32 master weights.
 * Raw network outputs are 48 floats per point: lobe j uses
 * [5j .. 5j+4] = mean (unnormalised xyz), log kappa, weight logit. */
typedef struct {
  int32_t t_log2;          /* table entries per level = 2^t_log2 (0: 19) */
  float learning_rate;     /* Adam step size (0: 5e-3) */
  float guide_fraction;    /* one-sample MIS: probability of sampling the guide (0: 0.5) */
  uint32_t seed;           /* parameter initialisation */
  float bounds_lo[3], bounds_hi[3];   /* scene box mapped to the unit cube */
  int32_t train_batch;     /* training samples per batch (0: 2^18) */
  int32_t record_paths;    /* paths per batch that record training samples (0: all; at most
                              2^25 / guide_depth, a hashed choice of the batch's paths) */
  int32_t train_steps;     /* optimisation steps per batch, each on an equal slice of its records
                              (0: one per 2^22 paths of a guided batch, 1 to 64; guided batches hold at
                              most 2^26 paths, or one job's pixels if more) */
  int32_t warmup_steps;    /* the guide's MIS probability ramps from 0 to guide_fraction over this many steps */
  int32_t guide_depth;     /* guide the bounces below this depth (0: 1, the camera vertex only: the best
                              equal-time trade-off measured on the caustic room); deeper vertices sample the BSDF */
} xnh_guide_config;

/* Turns guided rendering on for this context with fresh parameters. From
 * then on every render samples exact-Lambertian surfaces by one-sample MIS
 * of the guide (probability guide_fraction) and the BSDF, and after every
 * batch trains the network on up to train_batch of the guided vertices of
 * record_paths paths of that batch, drawn in proportion to their loss
 * weight (all on the device). At depth 0 a pixel's samples within a batch
 * share the mixture of its first sample. */
int xnh_guide_enable(xnh_ctx* ctx, const xnh_guide_config* cfg);
int xnh_guide_disable(xnh_ctx* ctx);
/* 1: train after every batch (default); 0: render with the network frozen. */
int xnh_guide_set_training(xnh_ctx* ctx, int32_t on);
/* FP32 master parameters: grid [16][2^t_log2][2], weights 48x64, 64x64,
 * 64x64, 64x48 (row-major [in][out], concatenated). */
int xnh_guide_set_params(xnh_ctx* ctx, const float* grid, const float* weights);
int xnh_guide_get_params(xnh_ctx* ctx, float* grid, float* weights);
int xnh_guide_eval(xnh_ctx* ctx, int32_t n, const float* pos, const float* nrm, float* raw_out);
/* Mixture queries: with u (3 per point) non-null, draws a direction per
 * point from its mixture into dir; then writes the mixture pdf (per
 * steradian) at dir into pdf. With u null, dir is an input. */
int xnh_guide_query(xnh_ctx* ctx, int32_t n, const float* pos, const float* nrm, const float* u, float* dir,
                    float* pdf);
/* One optimisation step on a host batch: loss -lambda_i log q(dir_i) with
 * lambda scaled to mean 1. With apply == 0 only gradients are computed and,
 * when grid_grad / w_grad are non-null, returned (sums over the batch of the
 * scaled loss); with apply != 0 Adam updates the parameters. */
int xnh_guide_train(xnh_ctx* ctx, int32_t n, const float* pos, const float* nrm, const float* dir, const float* lam,
                    int32_t apply, float* loss_out, float* grid_grad, float* w_grad);

/* ---- Tuning (no counterpart in the reference, whose CUDA backend fixes
 * 256-thread blocks at cuda_renderer.cu:29) ----
 * Workgroup size of the traversal kernels (64..256, multiple of 64), LDS
 * traversal-stack entries per lane (1..128; deeper entries spill to HBM) and
 * samples per pixel per batch (sets the queue capacity; 0 = auto, about
 * 2^28 paths, at most 256 samples: 292 B per path, about 75 GB). Changing them reallocates the queues and clears the running
 * accumulation. If the new queues do not fit in device memory, the previous
 * tuning is restored and the call returns XNH_E_ALLOC; if that too fails,
 * renders return XNH_E_ALLOC until a later xnh_set_tuning succeeds. */
typedef struct {
  int32_t block_size, lds_stack, spp_per_launch;
  float ms_per_spp;       /* device time of the winner per sample per pixel */
  int32_t candidates;     /* configurations timed */
} xnh_tune_result;

int xnh_set_tuning(xnh_ctx* ctx, int32_t block_size, int32_t lds_stack, int32_t spp_per_launch);
int xnh_get_tuning(const xnh_ctx* ctx, int32_t* block_size, int32_t* lds_stack, int32_t* spp_per_launch);
/* Times candidate configurations on the uploaded scene through `camera`
 * and keeps the fastest, in two stages: samples per batch (half, the
 * configured value and double) at the configured workgroup size and LDS
 * depth, then workgroup size (64, 128, 256) x LDS depth (8, 12, 16, 24, 32)
 * at the chosen batch. Each candidate renders one untimed batch, then
 * max(probe_spp, 2 x the stage's largest batch) timed samples. With guiding
 * on, the batch size stays fixed and the network is not trained during the
 * probes. Clears the accumulation. */
int xnh_autotune(xnh_ctx* ctx, const xnh_camera* camera, int32_t probe_spp, xnh_tune_result* out);

/* ---- output stage (main.cpp:25-52, math/image.h:11-182) ----------------
 * Both calls take a host framebuffer of W*H*3 floats in the layout the
 * reference's save_image receives (the swapchain buffer: rows top first),
 * or rgb = NULL for this context's running average in that layout (what
 * render_frame writes; W and H must then be the context's). */
typedef struct {           /* NoiseStats (math/image.h:11-31) */
  float mean_squared_score, rms_score, mean_absolute, median_absolute, p95_absolute, max_absolute;
  float avg_luminance, edge_threshold, epsilon;
  int32_t valid_pixel_count, total_interior_pixel_count;
  float valid_pixel_ratio;
} xnh_noise_stats;
/* ImageNoiseEstimator::computeResidualNoiseStats (image.h:35-118) on the
 * device. Images of width or height <= 2 give the all-zero result, as there.
 * At most 2^26 pixels. */
int xnh_image_noise_stats(xnh_ctx* ctx, const float* rgb, int32_t width, int32_t height, xnh_noise_stats* out);
/* xnh_image_noise_stats on a framebuffer already in device memory, with no
 * allocation: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). rgb_dev: W*H*3 floats in device memory,
 * the layout above, inside one device allocation (checked). The scratch is
 * one static 32 MiB array per GPU, so at most 2^22 pixels (2048 x 2048);
 * larger images return XNH_E_INVALID. The device address of that array is
 * looked up once per device by xnh_create. Calls share the array: calls
 * that overlap on one GPU (from two contexts or threads) race. The results
 * equal xnh_image_noise_stats's on the same image. */
int xnh_image_noise_stats_noalloc(xnh_ctx* ctx, const float* rgb_dev, int32_t width, int32_t height,
                                  xnh_noise_stats* out);
/* The 8-bit pixels save_image passes to stbi_write_png (main.cpp:26-31):
 * (uint8)(255.99 * clamp(c, 0, 1)^(1/2.2)), W*H*3 bytes in the file's row
 * order (the writer flips the framebuffer vertically, so file row r is
 * framebuffer row H-1-r). */
int xnh_image_rgb8(xnh_ctx* ctx, const float* rgb, int32_t width, int32_t height, uint8_t* out);

/* Copies the raw per-pixel sums (W*H*3) and returns the sample count. */
int xnh_read_accum(xnh_ctx* ctx, float* sum_out, int32_t* spp_out);
int xnh_reset(xnh_ctx* ctx);
int xnh_spp(const xnh_ctx* ctx);
int xnh_get_stats(const xnh_ctx* ctx, xnh_stats* out);
/* Profiling of renders, a mask of the flags below (0 = off, the default):
 * XNH_PROFILE_TIMING records HIP events around every traversal and shading
 * kernel (stats closest_ms / shade_ms / shadow_ms; the batch sequence is then
 * launched directly instead of as a graph); XNH_PROFILE_WORK runs the
 * counting variants of the traversal kernels (stats node_tests, interior,
 * tri_tests, inst_visits and their shadow counterparts). Ray and path counts
 * are always kept. */
#define XNH_PROFILE_TIMING 1
#define XNH_PROFILE_WORK 2
int xnh_set_profiling(xnh_ctx* ctx, int32_t flags);

/* ---- ray-batch parity API ----------------------------------------------- */
/* Scene::intersect / Scene::intersects (scene/scene.h:126-131) for a batch
 * of host rays; results copied back to host. */
int xnh_trace_closest(xnh_ctx* ctx, const xnh_ray* rays, int32_t n, xnh_hit* hits);
int xnh_trace_any(xnh_ctx* ctx, const xnh_ray* rays, int32_t n, uint8_t* occluded);

/* ---- host scene builder -------------------------------------------------
 * For callers without a reference Scene: the reference's scene model and
 * Scene::build_acceleration (scene/scene.h:79-121) with the reference's
 * binned-SAH BLAS/TLAS builders (geometry/blas.cpp:97-203,
 * geometry/tlas.cpp:40-145), producing the same trees node for node.
 * Material parameters: 16 floats (base rgb, roughness, metallic, ior,
 * transmission, subsurface, subsurface rgb, anisotropy, emission rgb, pad);
 * prepared as material_prepare does (material/material.h:55-75). */
xnh_scene* xnh_scene_new(void);
void xnh_scene_free(xnh_scene* s);
int xnh_scene_add_material(xnh_scene* s, const float* params16);
/* pos/nrm interleaved xyz; nrm NULL -> area-weighted smooth normals
 * (geometry/mesh.cpp:12-34). */
int xnh_scene_add_mesh(xnh_scene* s, int32_t nv, const float* pos, const float* nrm, int32_t nt,
                       const int32_t* indices, const int32_t* mat_ids);
int xnh_scene_add_light(xnh_scene* s, int32_t mesh, int32_t tri, const float* emission3, float area);
int xnh_scene_add_dir_light(xnh_scene* s, const float* dir3, const float* color3, float intensity);
int xnh_scene_set_sky(xnh_scene* s, const float* zenith3, const float* horizon3, const float* ground3,
                      float intensity, float sharpness);
/* Analytic primitive (xnh_prim types); returns its index (the hit record's
 * prim_id). Plane and disk normals are used as given (the scene parser
 * normalises them, scene_file.cpp:197-222). All primitives go into one
 * PrimBLAS, added to the TLAS as one identity instance after the instances
 * of the meshes (Scene::build_acceleration, scene/scene.h:97-110). */
int xnh_scene_add_primitive(xnh_scene* s, int32_t type, const float* params, int32_t mat_id);
/* TRS instance (geometry/transform.h:105-148); none -> identity per mesh. */
int xnh_scene_add_instance(xnh_scene* s, int32_t mesh, const float* t3, float scale, const float* rot_deg3,
                           int32_t instance_id);
int xnh_scene_build(xnh_scene* s, int32_t nthreads);
/* The same scene with every mesh's BLAS built on ctx's device by
 * xnh_build_blas (a linear BVH, not the reference's binned-SAH tree) and
 * the TLAS and primitives' BVH on the host as xnh_scene_build makes them. */
int xnh_scene_build_device(xnh_scene* s, xnh_ctx* ctx);

/* ---- BLAS on the device (no counterpart in the reference, which builds
 * on the CPU: blas.cpp:97-203) ------------------------------------------
 * A linear BVH (Morton order of the triangle box centres, Karras's radix
 * tree, boxes bottom-up) in the reference's node format: 2n-1 nodes, one
 * triangle per leaf, children adjacent and after their parent. The tree
 * differs from BLAS::build's, so hits agree with the reference's except
 * where two triangles tie at equal t (shared edges), and images agree
 * statistically. mesh: vertices and indices (its nodes are ignored);
 * nodes_out: 2n-1 nodes, prims_out: n slots, *num_nodes_out: 2n-1 (0 for
 * an empty mesh). */
int xnh_build_blas(xnh_ctx* ctx, const xnh_mesh* mesh, xnh_bvh_node* nodes_out, uint32_t* prims_out,
                   int32_t* num_nodes_out);
/* Recomputes the boxes of mesh's BLAS (any tree in the node format, e.g.
 * BLAS::build's or xnh_build_blas's) for its current vertices, keeping the
 * topology: leaves bound their triangles, interior nodes their children.
 * nodes_out: mesh->num_nodes nodes. */
int xnh_refit_blas(xnh_ctx* ctx, const xnh_mesh* mesh, xnh_bvh_node* nodes_out);
/* View of the built scene; valid until the scene is modified or freed. */
int xnh_scene_flat(const xnh_scene* s, xnh_flat_scene* out);
/* Camera::look_at (camera/camera.h:15-32). */
int xnh_camera_look_at(const float* eye3, const float* target3, const float* up3, float fov_deg, float aspect,
                       xnh_camera* out);

/* ---- scene files --------------------------------------------------------
 * load_scene (scene/scene_file.cpp:49-240): a .xenon scene file with its
 * .mat materials (material/material.cpp:31-152) and OBJ meshes
 * (geometry/mesh.cpp:42-205) into a built host scene plus the file's camera
 * and render settings. Relative paths resolve as the reference resolves them:
 * a matfile name as "materials/<name>" under the working directory, else
 * beside the scene file; a mesh path as given.
 * Analytic primitives (sphere, box, disk, quad, plane) are loaded as the
 * parser does (an unknown material name: material 0, counted in
 * unknown_materials). Files that cannot be opened are counted in
 * missing_files and skipped (a .mat falls back to the default material), as
 * the reference skips them.
 * One directive is this backend's own (the reference's parser skips it):
 *   instance <mesh> <tx> <ty> <tz> <scale> <rx> <ry> <rz> [id]
 * places mesh number <mesh> (file order) with the TRS of
 * xnh_scene_add_instance; id defaults to the instance's index. With any
 * instance line the TLAS holds exactly those instances; without, one
 * identity instance per mesh, as load_scene builds. Either way the
 * primitives, if any, follow as one identity instance (mesh
 * XNH_PRIM_GEOMETRY, id = the number of instances before it), as
 * Scene::build_acceleration adds them (scene/scene.h:97-110). A bad instance line
 * (unknown mesh, missing fields) fails the load with XNH_E_INVALID. */
typedef struct {
  int32_t width, height, samples;       /* SceneConfig (scene/scene_file.h:11-19) */
  int32_t min_bounces, max_bounces;
  int32_t meshes;                       /* meshes loaded */
  int32_t primitives;                   /* analytic primitives loaded */
  int32_t missing_files;                /* .mat / .obj files that did not open */
  int32_t instances;                    /* instance lines (0: one identity instance per mesh) */
  int32_t unknown_materials;            /* primitive lines naming no known material */
  float eye[3], target[3], up[3], fov;  /* the camera directive, for other aspect ratios */
} xnh_file_info;
int xnh_scene_load_file(const char* path, xnh_scene** out, xnh_camera* camera, xnh_file_info* info);
/* load_material (material/material.h:81-85) without the cache: the 16
 * material parameters above from a .mat file, defaults for absent keys;
 * name (NUL-terminated, truncated to name_cap) optional. XNH_E_INVALID when
 * the file does not open (params16 then holds the defaults). */
int xnh_material_load(const char* path, float* params16, char* name, int32_t name_cap);

#ifdef __cplusplus
}
#endif
#endif
