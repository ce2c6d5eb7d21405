/* oracle_api.h — C API of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Two libraries export exactly these symbols:
 *   oracle/lib/liborc_strict.so   the restatement (oracle/src), strict IEEE
 *   oracle/lib/liborc_fast.so     the same sources with the reference's
 *                                 Release flags (CPU-baseline timing)
 *   oracle/_ref/libxenon_ref*.so  the reference's own render core
 *                                 (compiled from its sources in place) behind
 *                                 oracle/ref_driver.cpp
 * so a test can run the same call sequence against the restatement and the
 * reference and compare bytes. Only tests/ and bench.py's CPU-baseline leg
 * load these libraries.
 *
 * Material parameters are 16 floats: base rgb, roughness, metallic, ior,
 * transmission, subsurface, subsurface rgb, anisotropy, emission rgb, pad.
 * Rays are 8 floats: origin, direction, tmin, tmax.
 */
#ifndef XENON_ORACLE_API_H
#define XENON_ORACLE_API_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

const char* orc_impl_name(void);
void* orc_scene_new(void);
void orc_scene_free(void* scene);
int orc_scene_add_material(void* scene, const float* params16);
void orc_material_derived(void* scene, int index, float* out11);
int orc_scene_add_mesh(void* scene, int nv, const float* pos, const float* nrm_or_null, int nt, const int* idx,
                       const int* mat_ids);
void orc_mesh_normals(void* scene, int mesh, float* out);
int orc_scene_add_light(void* scene, int mesh, int tri, const float* emission3, float area);
int orc_scene_add_dir_light(void* scene, const float* dir3, const float* color3, float intensity);
void orc_scene_set_sky(void* scene, const float* zenith3, const float* horizon3, const float* ground3,
                       float intensity, float sharpness);
int orc_scene_add_instance(void* scene, int mesh, const float* t3, float scale, const float* rot_deg3, int id);
int orc_scene_build(void* scene);
void orc_scene_set_camera(void* scene, const float* eye3, const float* target3, const float* up3, float fov_deg,
                          float aspect);
void orc_camera_get(void* scene, float* out12);
int orc_blas_node_count(void* scene, int mesh);
void orc_blas_export(void* scene, int mesh, void* nodes32, uint32_t* slots);
int orc_tlas_node_count(void* scene);
int orc_instance_count(void* scene);
void orc_tlas_export(void* scene, void* nodes32, uint32_t* order);
void orc_instance_export(void* scene, float* xf24, int* ids, int* meshes);
void orc_trace_closest(void* scene, int n, const float* rays8, float* hit12, int* hit4, uint64_t* counters4);
void orc_trace_any(void* scene, int n, const float* rays8, uint8_t* occluded, uint64_t* counters4);
int orc_render(void* scene, int W, int H, int spp_first, int spp_count, int min_bounces, int max_bounces,
               int nthreads, float* accum, uint64_t* stats3);
void orc_path(void* scene, int W, int H, int x, int y, int spp, int min_bounces, int max_bounces, float* out3);
int orc_path_rays(void* scene, int W, int H, int x, int y, int spp, int min_bounces, int max_bounces, int max_rays,
                  float* out11);
void orc_pcg_stream(uint64_t seed, uint64_t seq, int n, uint32_t* out);
void orc_bsdf_sample(const float* params16, const float* wo3, uint64_t seed, int n, float* out9);
void orc_bsdf_eval(const float* params16, int n, const float* wo3, const float* wi3, float* out4);

/* NoiseStats of a W x H x 3 framebuffer in the layout save_image receives
 * (math/image.h:35-118). out9: meanSquaredScore, rmsScore, meanAbsolute,
 * medianAbsolute, p95Absolute, maxAbsolute, avgLuminance, edgeThreshold,
 * epsilon; out3: validPixelCount, totalInteriorPixelCount, and
 * validPixelRatio as float bits. */
void orc_noise_stats(const float* rgb, int w, int h, float* out9, int* out3);
/* The 8-bit pixels save_image hands to the PNG writer (main.cpp:26-31),
 * rows in file order (the writer flips the framebuffer vertically). */
void orc_rgb8(const float* rgb, int w, int h, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif
