// hip_renderer.h — the reference-side binding of the HIP backend: the class a
// maintainer adds to Xenon as src/hip/hip_renderer.h, beside CudaRenderer
// (src/cuda/cuda_renderer.cuh:14-55), with the same three-method contract.
// It is compiled against the reference's own headers and linked with
// libxenon_hip.so; nothing in it is device code.
#pragma once
#include <vector>

#include "xenon_hip.h"

namespace xn {
struct Scene;
class Camera;
class TripleSwapchain;

// Everything an xnh_flat_scene points to that the Scene does not already
// hold in the library's layout. Mesh vertex arrays, BLAS nodes and leaf
// slots and the TLAS are referenced in place (their layouts are the
// library's: blas.h:21-28, tlas.h:91-96, mesh.h:21-31); the rest is copied
// here. `flat` is valid while this object and the Scene live.
struct HipFlatScene {
  std::vector<xnh_mesh> meshes;
  std::vector<xnh_instance> instances;
  std::vector<xnh_material> materials;
  std::vector<xnh_light> lights;
  std::vector<xnh_dir_light> suns;
  std::vector<xnh_prim> prims;
  xnh_scene* prim_scene = nullptr;   // the library's PrimBLAS over `prims`
  xnh_flat_scene flat{};

  HipFlatScene() = default;
  HipFlatScene(const HipFlatScene&) = delete;
  HipFlatScene& operator=(const HipFlatScene&) = delete;
  ~HipFlatScene();
};

// Flattens a Scene after build_acceleration (scene/scene.h:79-121) into
// `out`. Returns 0 or a negative XNH_E_* code.
int hip_flatten_scene(const Scene& scene, HipFlatScene& out);

class HipRenderer {
 public:
  HipRenderer(int width, int height, int max_bounces, int min_bounces);
  ~HipRenderer();
  HipRenderer(const HipRenderer&) = delete;
  HipRenderer& operator=(const HipRenderer&) = delete;

  void upload_scene(const Scene& scene);
  // Adds one sample per pixel and writes the running average into the
  // swapchain's write buffer, rows flipped as render/cpu_tiled.cpp:228-229.
  void render_frame(const Scene& scene, const Camera& camera, TripleSwapchain& chain);
  int spp() const { return xnh_spp(ctx_); }

 private:
  xnh_ctx* ctx_ = nullptr;
  int w_, h_;
  std::vector<float> img_;
};

}  // namespace xn
