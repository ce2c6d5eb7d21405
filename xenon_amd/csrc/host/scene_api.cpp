// scene_api.cpp — C ABI of the host scene builder (include/xenon_hip.h,
// "host scene builder" section): the reference's Scene model plus
// Scene::build_acceleration (scene/scene.h:79-121), for callers that do not
// hold a reference Scene. Compiled with g++ -ffp-contract=off.
#include <cmath>
#include <new>
#include <vector>

#include "scene.h"
#include "xenon_hip.h"

using namespace xnh;

struct xnh_scene {
  HostScene s;
};

extern "C" {

xnh_scene* xnh_scene_new(void) { return new (std::nothrow) xnh_scene(); }
void xnh_scene_free(xnh_scene* s) { delete s; }

int xnh_scene_add_material(xnh_scene* s, const float* p) {
  if (!s || !p) return XNH_E_INVALID;
  s->s.materials.push_back(prepare_material(p));
  s->s.built = false;
  return static_cast<int>(s->s.materials.size()) - 1;
}

int xnh_scene_add_mesh(xnh_scene* s, int32_t nv, const float* pos, const float* nrm, int32_t nt,
                       const int32_t* indices, const int32_t* mat_ids) {
  if (!s || nv < 0 || nt < 0 || (nv > 0 && !pos) || (nt > 0 && (!indices || !mat_ids))) return XNH_E_INVALID;
  HostMesh m;
  m.vx.resize(nv);
  m.vy.resize(nv);
  m.vz.resize(nv);
  for (int i = 0; i < nv; ++i) {
    m.vx[i] = pos[3 * i];
    m.vy[i] = pos[3 * i + 1];
    m.vz[i] = pos[3 * i + 2];
  }
  m.indices.assign(indices, indices + 3 * static_cast<size_t>(nt));
  m.mat_ids.assign(mat_ids, mat_ids + nt);
  for (int32_t v : m.indices)
    if (v < 0 || v >= nv) return XNH_E_INVALID;
  // Non-finite coordinates would index the SAH bins out of range (a NaN
  // centroid has no bin); the reference has no guard, this builder refuses.
  for (int i = 0; i < nv; ++i)
    if (!std::isfinite(m.vx[i]) || !std::isfinite(m.vy[i]) || !std::isfinite(m.vz[i])) return XNH_E_INVALID;
  if (nrm) {
    m.nx.resize(nv);
    m.ny.resize(nv);
    m.nz.resize(nv);
    for (int i = 0; i < nv; ++i) {
      m.nx[i] = nrm[3 * i];
      m.ny[i] = nrm[3 * i + 1];
      m.nz[i] = nrm[3 * i + 2];
    }
  } else {
    smooth_normals(m);
  }
  s->s.meshes.push_back(std::move(m));
  s->s.built = false;
  return static_cast<int>(s->s.meshes.size()) - 1;
}

int xnh_scene_add_light(xnh_scene* s, int32_t mesh, int32_t tri, const float* e, float area) {
  if (!s || !e || mesh < 0 || tri < 0) return XNH_E_INVALID;
  xnh_light l{};
  l.mesh_id = static_cast<uint32_t>(mesh);
  l.tri_idx = static_cast<uint32_t>(tri);
  l.emission[0] = e[0];
  l.emission[1] = e[1];
  l.emission[2] = e[2];
  l.area = area;
  s->s.lights.push_back(l);
  return static_cast<int>(s->s.lights.size()) - 1;
}

// Normalised as the scene parser does (scene_file.cpp:137-140).
int xnh_scene_add_dir_light(xnh_scene* s, const float* d, const float* color, float intensity) {
  if (!s || !d || !color) return XNH_E_INVALID;
  xnh_dir_light l{};
  const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (len > 1e-6f) {
    const float r = 1.f / len;
    l.direction[0] = d[0] * r;
    l.direction[1] = d[1] * r;
    l.direction[2] = d[2] * r;
  } else {
    l.direction[0] = 0.f;
    l.direction[1] = 1.f;
    l.direction[2] = 0.f;
  }
  for (int k = 0; k < 3; ++k) l.emission[k] = color[k] * intensity;
  s->s.suns.push_back(l);
  return static_cast<int>(s->s.suns.size()) - 1;
}

int xnh_scene_set_sky(xnh_scene* s, const float* z, const float* h, const float* g, float intensity, float sharp) {
  if (!s || !z || !h || !g) return XNH_E_INVALID;
  for (int k = 0; k < 3; ++k) {
    s->s.sky.zenith[k] = z[k];
    s->s.sky.horizon[k] = h[k];
    s->s.sky.ground[k] = g[k];
  }
  s->s.sky.intensity = intensity;
  s->s.sky.sharpness = sharp;
  return XNH_OK;
}

int xnh_scene_add_primitive(xnh_scene* s, int32_t type, const float* params, int32_t mat_id) {
  if (!s || !params || type < XNH_PRIM_SPHERE || type > XNH_PRIM_QUAD || mat_id < 0) return XNH_E_INVALID;
  static const int kParams[5] = {4, 4, 6, 7, 9};
  xnh_prim pr{};
  pr.type = type;
  pr.mat_id = mat_id;
  for (int k = 0; k < kParams[type]; ++k) {
    if (!std::isfinite(params[k])) return XNH_E_INVALID;
    pr.p[k] = params[k];
  }
  s->s.prims.push_back(pr);
  s->s.built = false;
  return static_cast<int>(s->s.prims.size()) - 1;
}

int xnh_scene_add_instance(xnh_scene* s, int32_t mesh, const float* t, float scale, const float* rot, int32_t id) {
  if (!s || !t || !rot || mesh < 0) return XNH_E_INVALID;
  PendingInstance p{};
  p.mesh = mesh;
  p.id = id;
  p.s = scale;
  for (int k = 0; k < 3; ++k) {
    p.t[k] = t[k];
    p.rot[k] = rot[k];
  }
  s->s.pending.push_back(p);
  s->s.built = false;
  return static_cast<int>(s->s.pending.size()) - 1;
}

namespace {

xnh_mesh geometry_view(const HostMesh& m) {
  xnh_mesh v{};
  v.num_vertices = static_cast<int32_t>(m.vx.size());
  v.num_triangles = m.tris();
  v.vx = m.vx.data();
  v.vy = m.vy.data();
  v.vz = m.vz.data();
  v.indices = m.indices.data();
  return v;
}

// The TLAS, the primitives' BVH and the flat views, once every BLAS exists.
int finish_build(HostScene& hs) {
  build_prims(hs);
  build_tlas(hs);
  hs.mesh_views.resize(hs.meshes.size());
  for (size_t i = 0; i < hs.meshes.size(); ++i) {
    HostMesh& m = hs.meshes[i];
    xnh_mesh& v = hs.mesh_views[i];
    v.num_vertices = static_cast<int32_t>(m.vx.size());
    v.num_triangles = m.tris();
    v.vx = m.vx.data();
    v.vy = m.vy.data();
    v.vz = m.vz.data();
    v.nx = m.nx.empty() ? nullptr : m.nx.data();
    v.ny = m.ny.empty() ? nullptr : m.ny.data();
    v.nz = m.nz.empty() ? nullptr : m.nz.data();
    v.indices = m.indices.data();
    v.mat_ids = m.mat_ids.data();
    v.num_nodes = static_cast<int32_t>(m.nodes.size());
    v.nodes = m.nodes.data();
    v.prims = m.prims.data();
  }
  hs.built = true;
  return XNH_OK;
}

}  // namespace

int xnh_scene_build(xnh_scene* s, int32_t nthreads) {
  (void)nthreads;
  if (!s) return XNH_E_INVALID;
  HostScene& hs = s->s;
  for (const auto& p : hs.pending)
    if (p.mesh >= static_cast<int32_t>(hs.meshes.size())) return XNH_E_INVALID;
  for (auto& m : hs.meshes) build_blas(m);
  return finish_build(hs);
}

int xnh_scene_build_device(xnh_scene* s, xnh_ctx* ctx) {
  if (!s || !ctx) return XNH_E_INVALID;
  HostScene& hs = s->s;
  for (const auto& p : hs.pending)
    if (p.mesh >= static_cast<int32_t>(hs.meshes.size())) return XNH_E_INVALID;
  for (auto& m : hs.meshes) {
    const xnh_mesh v = geometry_view(m);
    const size_t n = static_cast<size_t>(m.tris());
    m.nodes.assign(n > 0 ? 2 * n - 1 : 0, xnh_bvh_node{});
    m.prims.assign(n, 0u);
    int32_t nn = 0;
    const int rc = xnh_build_blas(ctx, &v, m.nodes.data(), m.prims.data(), &nn);
    if (rc != XNH_OK) return rc;
  }
  return finish_build(hs);
}

int xnh_scene_flat(const xnh_scene* s, xnh_flat_scene* out) {
  if (!s || !out || !s->s.built) return XNH_E_INVALID;
  const HostScene& hs = s->s;
  out->num_meshes = static_cast<int32_t>(hs.mesh_views.size());
  out->meshes = hs.mesh_views.data();
  out->num_instances = static_cast<int32_t>(hs.instances.size());
  out->instances = hs.instances.data();
  out->num_tlas_nodes = static_cast<int32_t>(hs.tlas_nodes.size());
  out->tlas_nodes = hs.tlas_nodes.data();
  out->tlas_inst_indices = hs.tlas_order.data();
  out->num_materials = static_cast<int32_t>(hs.materials.size());
  out->materials = hs.materials.data();
  out->num_lights = static_cast<int32_t>(hs.lights.size());
  out->lights = hs.lights.data();
  out->num_dir_lights = static_cast<int32_t>(hs.suns.size());
  out->dir_lights = hs.suns.data();
  out->sky = hs.sky;
  out->num_prims = static_cast<int32_t>(hs.prims.size());
  out->prims = hs.prims.data();
  out->num_prim_nodes = static_cast<int32_t>(hs.prim_nodes.size());
  out->prim_nodes = hs.prim_nodes.data();
  out->prim_order = hs.prim_order.data();
  return XNH_OK;
}

int xnh_camera_look_at(const float* eye, const float* target, const float* up, float fov, float aspect,
                       xnh_camera* out) {
  if (!eye || !target || !up || !out) return XNH_E_INVALID;
  *out = look_at(eye, target, up, fov, aspect);
  return XNH_OK;
}

}  // extern "C"
