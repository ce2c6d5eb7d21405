// hip_renderer.cpp — src/hip/hip_renderer.cpp of the reference-side binding
// (declared in hip_renderer.h). Reads the host Scene that build_acceleration
// produced through accessors the reference already has: BLAS::nodes() and
// prims() (geometry/blas.h:61-62), TLAS::nodes(), inst_indices() and
// instances() (geometry/tlas.h:118-120), the TriangleMesh SoA arrays
// (geometry/mesh.h:21-28) and Camera::get_* (camera/camera.h:40-43).
#include "hip_renderer.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <variant>

#include "camera/camera.h"
#include "render/swapchain.h"
#include "scene/scene.h"

namespace xn {

static_assert(sizeof(BLASNode) == sizeof(xnh_bvh_node) && sizeof(TLASNode) == sizeof(xnh_bvh_node),
              "BVH node layouts differ");
static_assert(sizeof(Mat4x3) == 48, "column-major 3x4 matrices");

namespace {

void check(xnh_ctx* c, int rc, const char* what) {
  if (rc < 0) {   // the CUDA backend aborts on errors too (cuda_utils.cuh:12-20)
    std::fprintf(stderr, "[HIP] %s failed: %d %s\n", what, rc, c ? xnh_last_error(c) : "");
    std::exit(1);
  }
}

void put3(float* dst, const Vec3& v) {
  dst[0] = v.x;
  dst[1] = v.y;
  dst[2] = v.z;
}

// One primitive in the parameter order of xnh_prim (xenon_hip.h); the
// variant's alternatives are numbered as XNH_PRIM_* (prim_blas.h:34).
xnh_prim to_prim(const PrimVariant& v) {
  xnh_prim q{};
  q.type = static_cast<int32_t>(v.index());
  q.mat_id = std::visit([](const auto& p) { return static_cast<int32_t>(p.mat_id); }, v);
  if (const auto* p = std::get_if<Sphere>(&v)) {
    put3(q.p, p->center);
    q.p[3] = p->radius;
  } else if (const auto* p = std::get_if<Plane>(&v)) {
    put3(q.p, p->normal);
    q.p[3] = p->offset;
  } else if (const auto* p = std::get_if<Box>(&v)) {
    put3(q.p, p->center);
    put3(q.p + 3, p->half);
  } else if (const auto* p = std::get_if<Disk>(&v)) {
    put3(q.p, p->center);
    put3(q.p + 3, p->normal);
    q.p[6] = p->radius;
  } else if (const auto* p = std::get_if<Quad>(&v)) {
    put3(q.p, p->origin);
    put3(q.p + 3, p->u_vec);
    put3(q.p + 6, p->v_vec);
  }
  return q;
}

}  // namespace

HipFlatScene::~HipFlatScene() { xnh_scene_free(prim_scene); }

int hip_flatten_scene(const Scene& s, HipFlatScene& o) {
  o.meshes.assign(s.meshes.size(), xnh_mesh{});
  for (size_t i = 0; i < s.meshes.size(); ++i) {
    const TriangleMesh& m = s.meshes[i];
    const BLAS& b = *s.blas_list[i];
    xnh_mesh& v = o.meshes[i];
    v.num_vertices = m.num_vertices();
    v.num_triangles = m.num_triangles();
    v.vx = m.vx.data();
    v.vy = m.vy.data();
    v.vz = m.vz.data();
    v.nx = m.nx.empty() ? nullptr : m.nx.data();
    v.ny = m.ny.empty() ? nullptr : m.ny.data();
    v.nz = m.nz.empty() ? nullptr : m.nz.data();
    v.indices = m.indices.data();
    v.mat_ids = m.mat_ids.data();
    v.num_nodes = static_cast<int32_t>(b.nodes().size());
    v.nodes = reinterpret_cast<const xnh_bvh_node*>(b.nodes().data());
    v.prims = b.prims().data();
  }
  o.instances.clear();
  for (const Instance& in : s.tlas.instances()) {
    xnh_instance x{};
    x.mesh = -2;
    if (in.geom.type == GeomHandle::PRIM_BLAS) {
      x.mesh = XNH_PRIM_GEOMETRY;
    } else {
      for (size_t k = 0; k < s.blas_list.size(); ++k)
        if (s.blas_list[k].get() == in.geom.mesh) x.mesh = static_cast<int32_t>(k);
    }
    if (x.mesh == -2) return XNH_E_INVALID;   // an instance of a BLAS the scene does not own
    x.instance_id = in.instance_id;
    std::memcpy(x.local_to_world, &in.xform.local_to_world, sizeof x.local_to_world);
    std::memcpy(x.world_to_local, &in.xform.world_to_local, sizeof x.world_to_local);
    o.instances.push_back(x);
  }
  o.materials.clear();
  for (const Material& m : s.materials) {
    xnh_material q{};
    put3(q.base_color, m.baseColor);
    q.metallic = m.metallic;
    q.ior = m.ior;
    q.alpha = m.alpha;
    q.alpha_x = m.alpha_x;
    q.alpha_y = m.alpha_y;
    put3(q.f0, m.F0);
    q.subsurface = m.subsurface;
    put3(q.subsurface_color, m.subsurfaceColor);
    q.flags = (m.isConductor ? XNH_MAT_CONDUCTOR : 0u) | (m.isTransmissive ? XNH_MAT_TRANSMISSIVE : 0u) |
              (m.hasSubsurface ? XNH_MAT_SUBSURFACE : 0u) | (m.isDelta ? XNH_MAT_DELTA : 0u) |
              (m.isEmissive ? XNH_MAT_EMISSIVE : 0u);
    o.materials.push_back(q);
  }
  o.lights.clear();
  for (const Light& l : s.lights) {
    xnh_light q{};
    q.mesh_id = l.mesh_id;
    q.tri_idx = l.tri_idx;
    put3(q.emission, l.emission);
    q.area = l.area;
    o.lights.push_back(q);
  }
  o.suns.clear();
  for (const DirectionalLight& d : s.dir_lights) {
    xnh_dir_light q{};
    put3(q.direction, d.direction);
    put3(q.emission, d.emission());
    o.suns.push_back(q);
  }
  // Primitives. PrimBLAS keeps its nodes and leaf order private
  // (prim_blas.h:97-100), so the library builds the same PrimBLAS from the
  // same list (12 bins, leaves of at most 2, prim_blas.cpp:34-140); the TLAS
  // instance over it is the reference's own.
  o.prims.clear();
  xnh_scene_free(o.prim_scene);
  o.prim_scene = nullptr;
  xnh_flat_scene pf{};
  if (!s.all_prims.empty()) {
    o.prim_scene = xnh_scene_new();
    if (!o.prim_scene) return XNH_E_ALLOC;
    for (const PrimVariant& v : s.all_prims.prims) {
      o.prims.push_back(to_prim(v));
      const xnh_prim& q = o.prims.back();
      const int rc = xnh_scene_add_primitive(o.prim_scene, q.type, q.p, q.mat_id);
      if (rc < 0) return rc;
    }
    int rc = xnh_scene_build(o.prim_scene, 1);
    if (rc == XNH_OK) rc = xnh_scene_flat(o.prim_scene, &pf);
    if (rc != XNH_OK) return rc;
  }
  xnh_flat_scene& f = o.flat;
  f = xnh_flat_scene{};
  f.num_meshes = static_cast<int32_t>(o.meshes.size());
  f.meshes = o.meshes.data();
  f.num_instances = static_cast<int32_t>(o.instances.size());
  f.instances = o.instances.data();
  f.num_tlas_nodes = static_cast<int32_t>(s.tlas.nodes().size());
  f.tlas_nodes = reinterpret_cast<const xnh_bvh_node*>(s.tlas.nodes().data());
  f.tlas_inst_indices = s.tlas.inst_indices().data();
  f.num_materials = static_cast<int32_t>(o.materials.size());
  f.materials = o.materials.data();
  f.num_lights = static_cast<int32_t>(o.lights.size());
  f.lights = o.lights.data();
  f.num_dir_lights = static_cast<int32_t>(o.suns.size());
  f.dir_lights = o.suns.data();
  put3(f.sky.zenith, s.sky.zenith_color);
  put3(f.sky.horizon, s.sky.horizon_color);
  put3(f.sky.ground, s.sky.ground_color);
  f.sky.sharpness = s.sky.horizon_sharpness;
  f.sky.intensity = s.sky.intensity;
  if (!o.prims.empty()) {
    f.num_prims = static_cast<int32_t>(o.prims.size());
    f.prims = o.prims.data();
    f.num_prim_nodes = pf.num_prim_nodes;
    f.prim_nodes = pf.prim_nodes;
    f.prim_order = pf.prim_order;
  }
  return XNH_OK;
}

HipRenderer::HipRenderer(int w, int h, int max_b, int min_b) : w_(w), h_(h), img_(static_cast<size_t>(w) * h * 3) {
  xnh_config cfg{};
  cfg.width = w;
  cfg.height = h;
  cfg.max_bounces = max_b;
  cfg.min_bounces = min_b;
  check(nullptr, xnh_create(&cfg, &ctx_), "xnh_create");
}

HipRenderer::~HipRenderer() { xnh_destroy(ctx_); }

void HipRenderer::upload_scene(const Scene& scene) {
  HipFlatScene flat;   // the library copies what it needs during the call
  check(ctx_, hip_flatten_scene(scene, flat), "flatten scene");
  check(ctx_, xnh_upload_scene(ctx_, &flat.flat), "xnh_upload_scene");
}

void HipRenderer::render_frame(const Scene&, const Camera& cam, TripleSwapchain& chain) {
  xnh_camera c{};
  put3(c.eye, cam.get_eye());
  put3(c.lower_left, cam.get_lower_left_corner());
  put3(c.horizontal, cam.get_horizontal());
  put3(c.vertical, cam.get_vertical());
  check(ctx_, xnh_render(ctx_, &c, xnh_spp(ctx_) + 1, 1, img_.data()), "xnh_render");
  float* out = chain.get_write_buffer();
  const size_t row = static_cast<size_t>(w_) * 3;
  for (int y = 0; y < h_; ++y)
    std::memcpy(out + static_cast<size_t>(h_ - 1 - y) * row, img_.data() + static_cast<size_t>(y) * row,
                sizeof(float) * row);
  chain.swap_writer();
}

}  // namespace xn
