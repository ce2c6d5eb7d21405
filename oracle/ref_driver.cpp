// ref_driver.cpp — headless harness over the reference's own render core.
//
// TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together with
// the reference's render-core sources, in place under /root/reference/src,
// into oracle/_ref/libxenon_ref.so (strict IEEE flags) and
// oracle/_ref/libxenon_ref_fast.so (the reference's Release flags). It
// exports the C API of oracle/oracle_api.h using the reference's own types
// (Scene, BLAS, TLAS, Camera, WavefrontRenderer, bsdf_*), so tests can
// compare the oracle restatement against the reference byte for byte.
// Nothing here is reference source; it only calls the reference's public
// interfaces.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "camera/camera.h"
#include "camera/sampler.h"
#include "material/bsdf.h"
#include "material/material.h"
#include "math/image.h"
#include "render/swapchain.h"
#include "render/wavefront.h"
#include "scene/scene.h"
#include "scene/scene_file.h"

#include "oracle_api.h"

namespace {
struct RefScene {
  xn::Scene scene;
  xn::Camera camera;
  struct PendingInst { int mesh; xn::Vec3 t; float s; xn::Vec3 rot; int id; };
  std::vector<PendingInst> pending;
  bool built = false;
};
RefScene* R(void* h) { return static_cast<RefScene*>(h); }

xn::Material material_from(const float* p) {
  xn::Material m;
  m.baseColor = xn::Vec3(p[0], p[1], p[2]);
  m.roughness = p[3];
  m.metallic = p[4];
  m.ior = p[5];
  m.transmission = p[6];
  m.subsurface = p[7];
  m.subsurfaceColor = xn::Vec3(p[8], p[9], p[10]);
  m.anisotropy = p[11];
  m.emission = xn::Vec3(p[12], p[13], p[14]);
  xn::material_prepare(m);
  return m;
}

void export_nodes(const void* src, size_t n, void* dst) { std::memcpy(dst, src, n * 32); }
}  // namespace

extern "C" {

const char* orc_impl_name(void) { return "reference"; }
void* orc_scene_new(void) { return new RefScene(); }
void orc_scene_free(void* h) { delete R(h); }

int orc_scene_add_material(void* h, const float* p) {
  R(h)->scene.materials.push_back(material_from(p));
  return static_cast<int>(R(h)->scene.materials.size()) - 1;
}

void orc_material_derived(void* h, int i, float* out) {
  const xn::Material& m = R(h)->scene.materials[i];
  out[0] = m.alpha; out[1] = m.alpha_x; out[2] = m.alpha_y;
  out[3] = m.F0.x; out[4] = m.F0.y; out[5] = m.F0.z;
  out[6] = m.isConductor; out[7] = m.isTransmissive; out[8] = m.hasSubsurface; out[9] = m.isDelta;
  out[10] = m.isEmissive;
}

// Fills a TriangleMesh the way load_obj does: vertices, triangles with
// per-triangle material ids, and smooth normals when none are supplied.
int orc_scene_add_mesh(void* h, int nv, const float* pos, const float* nrm, int nt, const int* idx, const int* mat) {
  xn::TriangleMesh m;
  for (int i = 0; i < nv; ++i) m.add_vertex(xn::Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
  for (int t = 0; t < nt; ++t) m.add_triangle(idx[3 * t], idx[3 * t + 1], idx[3 * t + 2], mat[t]);
  if (nrm) {
    for (int i = 0; i < nv; ++i) { m.nx[i] = nrm[3 * i]; m.ny[i] = nrm[3 * i + 1]; m.nz[i] = nrm[3 * i + 2]; }
  } else {
    m.compute_smooth_normals();
  }
  R(h)->scene.meshes.push_back(std::move(m));
  return static_cast<int>(R(h)->scene.meshes.size()) - 1;
}

void orc_mesh_normals(void* h, int mesh, float* out) {
  const xn::TriangleMesh& m = R(h)->scene.meshes[mesh];
  for (size_t i = 0; i < m.nx.size(); ++i) { out[3 * i] = m.nx[i]; out[3 * i + 1] = m.ny[i]; out[3 * i + 2] = m.nz[i]; }
}

int orc_scene_add_light(void* h, int mesh, int tri, const float* e, float area) {
  xn::Light lt{};
  lt.mesh_id = static_cast<uint32_t>(mesh);
  lt.tri_idx = static_cast<uint32_t>(tri);
  lt.emission = xn::Vec3(e[0], e[1], e[2]);
  lt.area = area;
  R(h)->scene.lights.push_back(lt);
  return static_cast<int>(R(h)->scene.lights.size()) - 1;
}

// Normalised with the same Vec3 operations the scene parser applies.
int orc_scene_add_dir_light(void* h, const float* d, const float* c, float intensity) {
  xn::DirectionalLight dl;
  dl.direction = xn::Vec3(d[0], d[1], d[2]);
  dl.color = xn::Vec3(c[0], c[1], c[2]);
  dl.intensity = intensity;
  const float len = dl.direction.length();
  if (len > 1e-6f) dl.direction = dl.direction / len;
  else dl.direction = xn::Vec3(0.f, 1.f, 0.f);
  R(h)->scene.dir_lights.push_back(dl);
  return static_cast<int>(R(h)->scene.dir_lights.size()) - 1;
}

void orc_scene_set_sky(void* h, const float* z, const float* hz, const float* g, float intensity, float sharp) {
  xn::Environment& s = R(h)->scene.sky;
  s.zenith_color = xn::Vec3(z[0], z[1], z[2]);
  s.horizon_color = xn::Vec3(hz[0], hz[1], hz[2]);
  s.ground_color = xn::Vec3(g[0], g[1], g[2]);
  s.intensity = intensity;
  s.horizon_sharpness = sharp;
}

int orc_scene_add_instance(void* h, int mesh, const float* t, float s, const float* rot, int id) {
  R(h)->pending.push_back({mesh, xn::Vec3(t[0], t[1], t[2]), s, xn::Vec3(rot[0], rot[1], rot[2]), id});
  return static_cast<int>(R(h)->pending.size()) - 1;
}

// Without explicit instances this is Scene::build_acceleration itself; with
// them, the same BLAS builds feed TLAS::build over the given instances.
int orc_scene_build(void* h) {
  RefScene& r = *R(h);
  if (r.pending.empty()) {
    r.scene.build_acceleration();
  } else {
    r.scene.blas_list.clear();
    for (auto& m : r.scene.meshes) {
      auto b = std::make_unique<xn::BLAS>();
      b->build(m);
      r.scene.blas_list.push_back(std::move(b));
    }
    std::vector<xn::Instance> inst;
    for (const auto& p : r.pending) {
      xn::Instance in;
      in.geom = xn::GeomHandle::from_mesh(r.scene.blas_list[p.mesh].get());
      in.xform = xn::AffineTransform::from_trs(p.t, p.s, p.rot);
      in.instance_id = p.id;
      in.rebuild_world_aabb();
      inst.push_back(std::move(in));
    }
    r.scene.tlas.build(std::move(inst));
  }
  r.built = true;
  return 0;
}

void orc_scene_set_camera(void* h, const float* e, const float* t, const float* u, float fov, float aspect) {
  R(h)->camera.look_at(xn::Vec3(e[0], e[1], e[2]), xn::Vec3(t[0], t[1], t[2]), xn::Vec3(u[0], u[1], u[2]), fov,
                       aspect);
}

void orc_camera_get(void* h, float* out) {
  const xn::Camera& c = R(h)->camera;
  const xn::Vec3 v[4] = {c.get_eye(), c.get_lower_left_corner(), c.get_horizontal(), c.get_vertical()};
  for (int i = 0; i < 4; ++i) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
}

int orc_blas_node_count(void* h, int mesh) { return static_cast<int>(R(h)->scene.blas_list[mesh]->nodes().size()); }
void orc_blas_export(void* h, int mesh, void* nodes, uint32_t* slots) {
  const xn::BLAS& b = *R(h)->scene.blas_list[mesh];
  export_nodes(b.nodes().data(), b.nodes().size(), nodes);
  std::memcpy(slots, b.prims().data(), b.prims().size() * sizeof(uint32_t));
}
int orc_tlas_node_count(void* h) { return static_cast<int>(R(h)->scene.tlas.nodes().size()); }
int orc_instance_count(void* h) { return static_cast<int>(R(h)->scene.tlas.instances().size()); }
void orc_tlas_export(void* h, void* nodes, uint32_t* order) {
  const xn::TLAS& t = R(h)->scene.tlas;
  export_nodes(t.nodes().data(), t.nodes().size(), nodes);
  std::memcpy(order, t.inst_indices().data(), t.inst_indices().size() * sizeof(uint32_t));
}
void orc_instance_export(void* h, float* xf24, int* ids, int* meshes) {
  const RefScene& r = *R(h);
  const auto& inst = r.scene.tlas.instances();
  for (size_t i = 0; i < inst.size(); ++i) {
    for (int c = 0; c < 4; ++c) {
      const xn::Vec3 a = inst[i].xform.local_to_world.cols[c];
      const xn::Vec3 b = inst[i].xform.world_to_local.cols[c];
      xf24[24 * i + 3 * c] = a.x; xf24[24 * i + 3 * c + 1] = a.y; xf24[24 * i + 3 * c + 2] = a.z;
      xf24[24 * i + 12 + 3 * c] = b.x; xf24[24 * i + 12 + 3 * c + 1] = b.y; xf24[24 * i + 12 + 3 * c + 2] = b.z;
    }
    ids[i] = inst[i].instance_id;
    int mesh = -1;
    for (size_t k = 0; k < r.scene.blas_list.size(); ++k)
      if (inst[i].geom.mesh == r.scene.blas_list[k].get()) mesh = static_cast<int>(k);
    meshes[i] = mesh;
  }
}

static xn::Ray load_ray(const float* r) {
  xn::Ray ray;
  ray.origin = xn::Vec3(r[0], r[1], r[2]);
  ray.dir = xn::Vec3(r[3], r[4], r[5]);
  ray.tmin = r[6];
  ray.tmax = r[7];
  return ray;
}

void orc_trace_closest(void* h, int n, const float* rays, float* hf, int* hi, uint64_t* counters) {
  const xn::Scene& sc = R(h)->scene;
  for (int i = 0; i < n; ++i) {
    xn::HitRecord rec;
    const bool ok = sc.intersect(load_ray(rays + 8 * i), rec);
    float* o = hf + 12 * i;
    int* oi = hi + 4 * i;
    if (!ok) {
      for (int k = 0; k < 12; ++k) o[k] = 0.f;
      o[0] = xn::kInfinity;
      oi[0] = -1; oi[1] = -1; oi[2] = -1; oi[3] = 0;
      continue;
    }
    const float v[12] = {rec.t, rec.u, rec.v, rec.pos.x, rec.pos.y, rec.pos.z, rec.normal.x, rec.normal.y,
                         rec.normal.z, rec.geo_normal.x, rec.geo_normal.y, rec.geo_normal.z};
    std::memcpy(o, v, sizeof(v));
    oi[0] = rec.instance_id; oi[1] = rec.prim_id; oi[2] = rec.mat_id; oi[3] = rec.front_face ? 1 : 0;
  }
  if (counters) for (int k = 0; k < 4; ++k) counters[k] = 0;
}

void orc_trace_any(void* h, int n, const float* rays, uint8_t* occ, uint64_t* counters) {
  const xn::Scene& sc = R(h)->scene;
  for (int i = 0; i < n; ++i) occ[i] = sc.intersects(load_ray(rays + 8 * i)) ? 1 : 0;
  if (counters) for (int k = 0; k < 4; ++k) counters[k] = 0;
}

// The reference renderer always starts at sample 1 and keeps its sums
// private; it only exposes the running average through the swapchain, with
// rows flipped. This entry therefore supports spp_first == 1 and returns
// the AVERAGE (not the sum) in pixel order. stats are not available.
int orc_render(void* h, int W, int H, int spp_first, int spp_count, int minb, int maxb, int nthreads, float* out,
               uint64_t* stats) {
  (void)nthreads;
  RefScene& r = *R(h);
  if (!r.built || spp_first != 1 || spp_count < 1) return -2;
  xn::WavefrontRenderer renderer(W, H);
  renderer.set_bounces(minb, maxb);
  xn::TripleSwapchain chain(W, H);
  for (int i = 0; i < spp_count; ++i) renderer.render_frame_tiled(r.scene, r.camera, chain);
  const float* img = chain.get_read_buffer();
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x)
      for (int c = 0; c < 3; ++c)
        out[3 * (static_cast<size_t>(y) * W + x) + c] = img[3 * (static_cast<size_t>(H - 1 - y) * W + x) + c];
  if (stats) { stats[0] = 0; stats[1] = 0; stats[2] = 0; }
  return 0;
}

void orc_path(void*, int, int, int, int, int, int, int, float* out3) {
  out3[0] = out3[1] = out3[2] = __builtin_nanf("");
}

void orc_pcg_stream(uint64_t seed, uint64_t seq, int n, uint32_t* out) {
  xn::PCGState p = xn::seed_pcg(seed, seq);
  for (int i = 0; i < n; ++i) out[i] = p.next_uint();
}

void orc_bsdf_sample(const float* p, const float* wo, uint64_t seed, int n, float* out) {
  const xn::Material m = material_from(p);
  xn::PCGState rng = xn::seed_pcg(seed);
  const xn::Vec3 w(wo[0], wo[1], wo[2]);
  for (int i = 0; i < n; ++i) {
    xn::BSDFSample s;
    const bool ok = xn::bsdf_sample(w, m, rng, s);
    float* o = out + 9 * i;
    o[0] = s.wi.x; o[1] = s.wi.y; o[2] = s.wi.z; o[3] = s.f.x; o[4] = s.f.y; o[5] = s.f.z;
    o[6] = s.pdf; o[7] = s.is_delta ? 1.f : 0.f; o[8] = ok ? 1.f : 0.f;
  }
}

void orc_bsdf_eval(const float* p, int n, const float* wo, const float* wi, float* out) {
  const xn::Material m = material_from(p);
  for (int i = 0; i < n; ++i) {
    const xn::Vec3 a(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    const xn::Vec3 b(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
    const xn::Vec3 f = xn::bsdf_eval_for_nee(a, b, m);
    out[4 * i] = f.x; out[4 * i + 1] = f.y; out[4 * i + 2] = f.z;
    out[4 * i + 3] = xn::bsdf_pdf_for_nee(a, b, m);
  }
}

// NoiseStats: the reference's own estimator (math/image.h, header only).
void orc_noise_stats(const float* rgb, int w, int h, float* o, int* oi) {
  const xn::NoiseStats s = xn::ImageNoiseEstimator::computeResidualNoiseStats(rgb, w, h);
  const float v[9] = {s.meanSquaredScore, s.rmsScore, s.meanAbsolute, s.medianAbsolute, s.p95Absolute,
                      s.maxAbsolute, s.avgLuminance, s.edgeThreshold, s.epsilon};
  std::memcpy(o, v, sizeof v);
  oi[0] = s.validPixelCount;
  oi[1] = s.totalInteriorPixelCount;
  std::memcpy(&oi[2], &s.validPixelRatio, 4);
}

// save_image's conversion (main.cpp:26-31) restated: main.cpp itself does
// not build without GLFW and stb. File row r is framebuffer row h-1-r.
void orc_rgb8(const float* rgb, int w, int h, uint8_t* out) {
  for (int r = 0; r < h; ++r)
    for (int k = 0; k < 3 * w; ++k) {
      const float c = std::pow(std::clamp(rgb[3 * static_cast<size_t>(h - 1 - r) * w + k], 0.f, 1.f), 1.f / 2.2f);
      out[3 * static_cast<size_t>(r) * w + k] = static_cast<uint8_t>(255.99f * c);
    }
}

// ---- analytic primitives (reference only; not part of oracle_api.h) ------
// One primitive into Scene::all_prims (geometry/prim_blas.h:54-61); p as
// xnh_prim (sphere c r | plane n offset | box c half | disk c n r | quad o u v).
int orc_scene_add_prim(void* h, int type, const float* p, int mat) {
  xn::PrimGroup& g = R(h)->scene.all_prims;
  switch (type) {
    case 0: { xn::Sphere s; s.center = xn::Vec3(p[0], p[1], p[2]); s.radius = p[3]; s.mat_id = mat; g.add(s); break; }
    case 1: { xn::Plane q; q.normal = xn::Vec3(p[0], p[1], p[2]); q.offset = p[3]; q.mat_id = mat; g.add(q); break; }
    case 2: { xn::Box b; b.center = xn::Vec3(p[0], p[1], p[2]); b.half = xn::Vec3(p[3], p[4], p[5]); b.mat_id = mat; g.add(b); break; }
    case 3: { xn::Disk d; d.center = xn::Vec3(p[0], p[1], p[2]); d.normal = xn::Vec3(p[3], p[4], p[5]); d.radius = p[6]; d.mat_id = mat; g.add(d); break; }
    case 4: { xn::Quad q; q.origin = xn::Vec3(p[0], p[1], p[2]); q.u_vec = xn::Vec3(p[3], p[4], p[5]); q.v_vec = xn::Vec3(p[6], p[7], p[8]); q.mat_id = mat; g.add(q); break; }
    default: return -1;
  }
  return g.size() - 1;
}

// Scene::all_prims back out: per primitive its type (PrimVariant index, the
// xnh_prim type numbering), mat_id and parameters in orc_scene_add_prim order.
void orc_prims_export(void* h, int* type_mat, float* p10) {
  const auto& P = R(h)->scene.all_prims.prims;
  for (size_t i = 0; i < P.size(); ++i) {
    float* o = p10 + 10 * i;
    for (int k = 0; k < 10; ++k) o[k] = 0.f;
    auto put = [&](int at, const xn::Vec3& v) { o[at] = v.x; o[at + 1] = v.y; o[at + 2] = v.z; };
    type_mat[2 * i] = static_cast<int>(P[i].index());
    type_mat[2 * i + 1] = xn::prim_mat_id(P[i]);
    if (auto* s = std::get_if<xn::Sphere>(&P[i])) { put(0, s->center); o[3] = s->radius; }
    else if (auto* q = std::get_if<xn::Plane>(&P[i])) { put(0, q->normal); o[3] = q->offset; }
    else if (auto* b = std::get_if<xn::Box>(&P[i])) { put(0, b->center); put(3, b->half); }
    else if (auto* d = std::get_if<xn::Disk>(&P[i])) { put(0, d->center); put(3, d->normal); o[6] = d->radius; }
    else if (auto* u = std::get_if<xn::Quad>(&P[i])) { put(0, u->origin); put(3, u->u_vec); put(6, u->v_vec); }
  }
}

// ---- scene files (reference only; not part of oracle_api.h) --------------
// load_scene (scene/scene_file.cpp:49-240) into a new handle; cfg5 receives
// SceneConfig's width, height, samples, min_bounces, max_bounces. NULL when
// the file does not open.
void* orc_ref_load_scene_file(const char* path, int* cfg5) {
  auto* r = new RefScene();
  xn::SceneConfig cfg;
  if (!xn::load_scene(path, r->scene, r->camera, cfg)) {
    delete r;
    return nullptr;
  }
  cfg5[0] = cfg.width; cfg5[1] = cfg.height; cfg5[2] = cfg.samples;
  cfg5[3] = cfg.min_bounces; cfg5[4] = cfg.max_bounces;
  r->built = true;
  return r;
}

// meshes, materials, lights, directional lights, analytic primitives
void orc_ref_scene_sizes(void* h, int* out5) {
  const xn::Scene& s = R(h)->scene;
  out5[0] = static_cast<int>(s.meshes.size());
  out5[1] = static_cast<int>(s.materials.size());
  out5[2] = static_cast<int>(s.lights.size());
  out5[3] = static_cast<int>(s.dir_lights.size());
  out5[4] = s.all_prims.size();
}

void orc_ref_mesh_size(void* h, int mesh, int* nv_nt) {
  const xn::TriangleMesh& m = R(h)->scene.meshes[mesh];
  nv_nt[0] = m.num_vertices();
  nv_nt[1] = m.num_triangles();
}

void orc_ref_mesh_export(void* h, int mesh, float* pos, float* nrm, int* idx, int* mat) {
  const xn::TriangleMesh& m = R(h)->scene.meshes[mesh];
  for (int i = 0; i < m.num_vertices(); ++i) {
    pos[3 * i] = m.vx[i]; pos[3 * i + 1] = m.vy[i]; pos[3 * i + 2] = m.vz[i];
    nrm[3 * i] = m.nx[i]; nrm[3 * i + 1] = m.ny[i]; nrm[3 * i + 2] = m.nz[i];
  }
  for (int t = 0; t < m.num_triangles(); ++t) {
    for (int k = 0; k < 3; ++k) idx[3 * t + k] = m.indices[3 * t + k];
    mat[t] = m.mat_ids[t];
  }
}

// authoring parameters in the 16-float order of material_from
void orc_ref_material_params(void* h, int i, float* p) {
  const xn::Material& m = R(h)->scene.materials[i];
  const float v[16] = {m.baseColor.x, m.baseColor.y, m.baseColor.z, m.roughness, m.metallic, m.ior,
                       m.transmission, m.subsurface, m.subsurfaceColor.x, m.subsurfaceColor.y,
                       m.subsurfaceColor.z, m.anisotropy, m.emission.x, m.emission.y, m.emission.z, 0.f};
  std::memcpy(p, v, sizeof v);
}

// per light: mesh, triangle (as float), emission rgb, area
void orc_ref_lights_export(void* h, float* out6) {
  const auto& L = R(h)->scene.lights;
  for (size_t i = 0; i < L.size(); ++i) {
    float* o = out6 + 6 * i;
    o[0] = static_cast<float>(L[i].mesh_id); o[1] = static_cast<float>(L[i].tri_idx);
    o[2] = L[i].emission.x; o[3] = L[i].emission.y; o[4] = L[i].emission.z; o[5] = L[i].area;
  }
}

// per directional light: direction, emission() = color * intensity
void orc_ref_suns_export(void* h, float* out6) {
  const auto& D = R(h)->scene.dir_lights;
  for (size_t i = 0; i < D.size(); ++i) {
    const xn::Vec3 e = D[i].emission();
    float* o = out6 + 6 * i;
    o[0] = D[i].direction.x; o[1] = D[i].direction.y; o[2] = D[i].direction.z;
    o[3] = e.x; o[4] = e.y; o[5] = e.z;
  }
}

// zenith, horizon, ground, sharpness, intensity
void orc_ref_sky_export(void* h, float* out11) {
  const xn::Environment& s = R(h)->scene.sky;
  const float v[11] = {s.zenith_color.x, s.zenith_color.y, s.zenith_color.z, s.horizon_color.x,
                       s.horizon_color.y, s.horizon_color.z, s.ground_color.x, s.ground_color.y,
                       s.ground_color.z, s.horizon_sharpness, s.intensity};
  std::memcpy(out11, v, sizeof v);
}

}  // extern "C"
