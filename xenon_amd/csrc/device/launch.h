// launch.h — host-callable launchers of the device kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../common/dev_types.h"
#include "../common/guide_net.h"
#include "xenon_hip.h"

namespace xnh {

// Persistent grid for the traversal kernels: CUs x resident blocks.
int trace_grid(int block, int lds_depth);
int num_cus();

hipError_t launch_raygen(hipStream_t s, const WaveBuffers& B, int max_paths);
// camera: depth 0 (the queue's rays are the chunks' camera rays)
hipError_t launch_closest(hipStream_t s, const WaveBuffers& B, int qi, int cursor, int grid, int block,
                          int lds_depth, bool count, bool prims, bool camera);
hipError_t launch_shade(hipStream_t s, const WaveBuffers& B, int qi, int depth, int grid);
hipError_t launch_shadow(hipStream_t s, const WaveBuffers& B, int cursor, int grid, int block, int lds_depth,
                         bool count, bool prims);
hipError_t launch_accumulate(hipStream_t s, const WaveBuffers& B, int max_pixels);
hipError_t launch_advance(hipStream_t s, uint32_t* counters);
hipError_t launch_average(hipStream_t s, const float4* accum, float* out, int npix, float inv_spp);
// Ray-batch kernels: rays in the xnh_ray layout (2 float4 per ray:
// {org.xyz, dir.x} {dir.yz, tmin, tmax}), hits in the xnh_hit layout.
hipError_t launch_api_closest(hipStream_t s, const DevScene& S, const float4* rays, float* hits, int n,
                              uint32_t* spill, int lds_depth, unsigned long long* stats, uint32_t* overflow_ctr);
hipError_t launch_api_any(hipStream_t s, const DevScene& S, const float4* rays, uint8_t* occ, int n, uint32_t* spill,
                          int lds_depth, unsigned long long* stats, uint32_t* overflow_ctr);

hipError_t launch_guide_eval(hipStream_t s, const GuideNet& net, int n, const float* pos, const float* nrm,
                             float* raw);
hipError_t launch_guide_query(hipStream_t s, const GuideNet& net, int n, const float* pos, const float* nrm,
                              const float* u, float* dir, float* pdf);
// Work-groups of one launch_guide_train call (each needs kWTotal floats of
// w_part): n records, or with device counts the largest slice of n.
int guide_train_blocks(int n, bool device_count, int nslices);
// Forward/backward over the records; grid gradients are added to grid_g,
// weight gradients go through w_part (per work-group) and are added to w_g.
hipError_t launch_guide_train(hipStream_t s, const GuideNet& net, int n, const float* pos, const float* nrm,
                              const float* dir, const float* lam, float lam_scale, const uint32_t* n_dev,
                              const float* lam_sum_dev, int slice, int nslices, float* grid_g, float* w_part,
                              float* w_g, float* loss_sum);
hipError_t launch_guide_adam(hipStream_t s, float* master, uint16_t* half_copy, float* grad, float* m, float* v,
                             size_t count, float lr, float b1, float b2, float eps, float bc1, float bc2,
                             float grad_scale, const uint32_t* n_dev, uint32_t cap, int nslices);
hipError_t launch_tile_weights(hipStream_t s, const uint16_t* w, uint16_t* wt);
// Guide mixtures for the hits of queue qi (encode, then MLP); feat: 6 uint4
// per queue entry of scratch.
hipError_t launch_guide_infer(hipStream_t s, const WaveBuffers& B, int qi, int depth, const GuideNet& net, uint4* feat,
                              int grid);
// Training batch of the batch just rendered: lambda of every record, then
// at most `cap` records drawn in proportion to lambda (target: expected
// count), written to pos/nrm/dir/lam; *n_out and *lam_sum must be zero.
hipError_t launch_guide_collect(hipStream_t s, const WaveBuffers& B, float* lam_all, float* lam_sum_all,
                                float target, uint32_t seed, float* pos, float* nrm, float* dir, float* lam,
                                float* lam_sum, uint32_t* n_out, uint32_t cap);
hipError_t launch_to_half(hipStream_t s, const float* src, uint16_t* dst, size_t count);

// Output stage (image.hip). img: W*H*3 floats in the framebuffer layout the
// PNG writer receives (rows top first).
hipError_t launch_img_from_accum(hipStream_t s, const float4* accum, float inv_spp, int w, int h, float* img);
hipError_t launch_noise_stats(hipStream_t s, const float* img, int w, int h, ImgScratch sc);
hipError_t launch_img_rgb8(hipStream_t s, const float* img, int w, int h, uint8_t* out);
// The static scratch array of image.hip: its device address is looked up once
// per device (with that device current) by img_static_scratch_init;
// img_static_scratch carves it into an ImgScratch.
hipError_t img_static_scratch_init(int device);
hipError_t img_static_scratch(int device, ImgScratch* sc);

// BLAS on the device (bvh.hip). build: n >= 2 triangles (vertex indices
// checked by the caller) -> 2n-1 nodes in the reference's node format and n
// leaf slots; refit: boxes of an existing tree in that format (structure
// checked by the caller). parent and flag: nn words of scratch each; err
// gets a nonzero bit mask if the tree came out malformed. build frees its
// own scratch and synchronises the stream.
hipError_t build_blas_device(hipStream_t s, const float* vx, const float* vy, const float* vz, const int32_t* idx,
                             int n, xnh_bvh_node* nodes, uint32_t* prims, uint32_t* err);
hipError_t refit_blas_device(hipStream_t s, xnh_bvh_node* nodes, int nn, const uint32_t* prims, int nprims,
                             const float* vx, const float* vy, const float* vz, const int32_t* idx, uint32_t* parent,
                             uint32_t* flag, uint32_t* err);

}  // namespace xnh
