// scene.h — host-side scene model and BVH construction of the HIP backend.
//
// This is the host half of the drop-in: the reference builds its BLAS/TLAS
// on the CPU (scene/scene.h:79-121) before CudaRenderer::upload_scene reads
// them. Callers that hold a reference Scene pass its arrays straight to
// xnh_upload_scene; this module serves callers that do not, and must
// therefore produce the reference's trees node for node, because traversal
// order (and with it the winner of equal-t hits) is a function of the tree.
#pragma once
#include <cstdint>
#include <vector>

#include "xenon_hip.h"

namespace xnh {

struct V3 {
  float x, y, z;
};

// Axis-aligned box with the reference's std::min/std::max growth semantics
// (geometry/aabb.h:22-29).
struct Aabb {
  V3 lo{kBig(), kBig(), kBig()};
  V3 hi{-kBig(), -kBig(), -kBig()};
  static float kBig();
  void add(const V3& p);
  void add(const Aabb& b);
  V3 center() const;
  V3 extent() const;
  float half_area2() const;  // the reference's surface_area(): 2(xy + yz + zx)
};

struct HostMesh {
  std::vector<float> vx, vy, vz, nx, ny, nz;
  std::vector<int32_t> indices, mat_ids;
  std::vector<xnh_bvh_node> nodes;
  std::vector<uint32_t> prims;
  int tris() const { return static_cast<int>(mat_ids.size()); }
};

struct PendingInstance {
  int32_t mesh, id;
  float t[3], s, rot[3];
};

struct HostScene {
  std::vector<HostMesh> meshes;
  std::vector<xnh_material> materials;
  std::vector<xnh_light> lights;
  std::vector<xnh_dir_light> suns;
  xnh_sky sky;
  std::vector<xnh_prim> prims;             // Scene::all_prims
  std::vector<xnh_bvh_node> prim_nodes;    // its PrimBLAS
  std::vector<uint32_t> prim_order;
  std::vector<PendingInstance> pending;
  std::vector<xnh_instance> instances;
  std::vector<xnh_bvh_node> tlas_nodes;
  std::vector<uint32_t> tlas_order;
  std::vector<xnh_mesh> mesh_views;
  bool built = false;
  HostScene();
};

xnh_material prepare_material(const float* p16);
void smooth_normals(HostMesh& m);
void build_blas(HostMesh& m);
void build_prims(HostScene& s);
void build_tlas(HostScene& s);
xnh_instance make_instance(int32_t mesh, int32_t id, const float t[3], float s, const float rot_deg[3]);
xnh_camera look_at(const float eye[3], const float target[3], const float up[3], float fov_deg, float aspect);

}  // namespace xnh
