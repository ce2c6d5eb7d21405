// check_binding.cpp — CPU test driver for the reference-side binding
// (xenon_amd/binding/hip_renderer.cpp). For each scene file named on the
// command line it builds the scene twice: the reference's own load_scene
// (scene/scene_file.cpp:49-240, whose build_acceleration builds the BVHs)
// flattened by hip_flatten_scene, and the library's xnh_scene_load_file
// flattened by xnh_scene_flat. Every array of the two xnh_flat_scene views
// must agree byte for byte. No GPU call is made. Prints one line per file;
// exits non-zero on the first difference.
#include <cstdio>
#include <cstring>
#include <string>

#include "camera/camera.h"
#include "hip_renderer.h"
#include "scene/scene.h"
#include "scene/scene_file.h"

namespace {

int g_fail = 0;

bool same(const void* a, const void* b, size_t bytes, const std::string& what) {
  if (bytes == 0) return true;
  if (!a || !b || std::memcmp(a, b, bytes) != 0) {
    std::printf("  MISMATCH %s (%zu bytes)\n", what.c_str(), bytes);
    g_fail = 1;
    return false;
  }
  return true;
}

bool same_count(long a, long b, const std::string& what) {
  if (a != b) {
    std::printf("  MISMATCH %s: %ld vs %ld\n", what.c_str(), a, b);
    g_fail = 1;
    return false;
  }
  return true;
}

void compare(const xnh_flat_scene& a, const xnh_flat_scene& b) {
  if (same_count(a.num_meshes, b.num_meshes, "mesh count")) {
    for (int i = 0; i < a.num_meshes; ++i) {
      const xnh_mesh &x = a.meshes[i], &y = b.meshes[i];
      const std::string m = "mesh " + std::to_string(i) + " ";
      if (!same_count(x.num_vertices, y.num_vertices, m + "vertices") ||
          !same_count(x.num_triangles, y.num_triangles, m + "triangles") ||
          !same_count(x.num_nodes, y.num_nodes, m + "nodes"))
        continue;
      const size_t nv = sizeof(float) * x.num_vertices, nt = x.num_triangles;
      same(x.vx, y.vx, nv, m + "vx");
      same(x.vy, y.vy, nv, m + "vy");
      same(x.vz, y.vz, nv, m + "vz");
      same_count(x.nx != nullptr, y.nx != nullptr, m + "normals present");
      if (x.nx && y.nx) {
        same(x.nx, y.nx, nv, m + "nx");
        same(x.ny, y.ny, nv, m + "ny");
        same(x.nz, y.nz, nv, m + "nz");
      }
      same(x.indices, y.indices, 3 * sizeof(int32_t) * nt, m + "indices");
      same(x.mat_ids, y.mat_ids, sizeof(int32_t) * nt, m + "mat_ids");
      same(x.nodes, y.nodes, sizeof(xnh_bvh_node) * x.num_nodes, m + "BLAS nodes");
      same(x.prims, y.prims, sizeof(uint32_t) * nt, m + "leaf slots");
    }
  }
  if (same_count(a.num_instances, b.num_instances, "instance count"))
    same(a.instances, b.instances, sizeof(xnh_instance) * a.num_instances, "instances");
  if (same_count(a.num_tlas_nodes, b.num_tlas_nodes, "TLAS node count")) {
    same(a.tlas_nodes, b.tlas_nodes, sizeof(xnh_bvh_node) * a.num_tlas_nodes, "TLAS nodes");
    same(a.tlas_inst_indices, b.tlas_inst_indices, sizeof(uint32_t) * a.num_instances, "TLAS instance order");
  }
  if (same_count(a.num_materials, b.num_materials, "material count"))
    same(a.materials, b.materials, sizeof(xnh_material) * a.num_materials, "materials");
  if (same_count(a.num_lights, b.num_lights, "light count"))
    same(a.lights, b.lights, sizeof(xnh_light) * a.num_lights, "lights");
  if (same_count(a.num_dir_lights, b.num_dir_lights, "sun count"))
    same(a.dir_lights, b.dir_lights, sizeof(xnh_dir_light) * a.num_dir_lights, "suns");
  same(&a.sky, &b.sky, sizeof(xnh_sky), "sky");
  if (same_count(a.num_prims, b.num_prims, "primitive count"))
    same(a.prims, b.prims, sizeof(xnh_prim) * a.num_prims, "primitives");
  if (same_count(a.num_prim_nodes, b.num_prim_nodes, "PrimBLAS node count")) {
    same(a.prim_nodes, b.prim_nodes, sizeof(xnh_bvh_node) * a.num_prim_nodes, "PrimBLAS nodes");
    same(a.prim_order, b.prim_order, sizeof(uint32_t) * a.num_prims, "PrimBLAS leaf order");
  }
}

}  // namespace

int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    const char* path = argv[i];
    xn::Scene scene;
    xn::Camera camera;
    xn::SceneConfig config;
    if (!xn::load_scene(path, scene, camera, config)) {
      std::printf("FAIL %s: the reference's load_scene failed\n", path);
      return 2;
    }
    xn::HipFlatScene ours;
    int rc = xn::hip_flatten_scene(scene, ours);
    if (rc != XNH_OK) {
      std::printf("FAIL %s: hip_flatten_scene returned %d\n", path, rc);
      return 2;
    }
    xnh_scene* lib_scene = nullptr;
    xnh_camera cam{};
    xnh_file_info info{};
    rc = xnh_scene_load_file(path, &lib_scene, &cam, &info);
    xnh_flat_scene theirs{};
    if (rc == XNH_OK) rc = xnh_scene_flat(lib_scene, &theirs);
    if (rc != XNH_OK) {
      std::printf("FAIL %s: xnh_scene_load_file/flat returned %d\n", path, rc);
      xnh_scene_free(lib_scene);
      return 2;
    }
    compare(ours.flat, theirs);
    const float c[12] = {camera.get_eye().x, camera.get_eye().y, camera.get_eye().z,
                         camera.get_lower_left_corner().x, camera.get_lower_left_corner().y,
                         camera.get_lower_left_corner().z, camera.get_horizontal().x, camera.get_horizontal().y,
                         camera.get_horizontal().z, camera.get_vertical().x, camera.get_vertical().y,
                         camera.get_vertical().z};
    same(c, &cam, sizeof c, "camera");
    std::printf("%s %s: %d meshes, %d instances, %d primitives, %d lights\n", g_fail ? "FAIL" : "OK", path,
                theirs.num_meshes, theirs.num_instances, theirs.num_prims, theirs.num_lights);
    xnh_scene_free(lib_scene);
    if (g_fail) return 1;
  }
  return 0;
}
