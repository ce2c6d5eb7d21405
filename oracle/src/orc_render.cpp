// orc_render.cpp — CPU oracle integrator and its C API.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP
// backend: tests/ and bench.py's CPU-baseline leg load it; the product
// library never does.
//
// The integrator restates render/cpu_tiled.cpp:13-237 (per-pixel
// unidirectional path tracing with area-light NEE, sun NEE, power-2 MIS,
// sky on miss and Russian roulette) over the scene container of
// scene/scene.h:52-161 and the sky of scene/environment.h:26-51.
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "../oracle_api.h"
#include "orc_bsdf.h"
#include "orc_geom.h"

namespace orc {

struct AreaLight {   // scene.h:19-24: emitter named by (mesh index, triangle)
  uint32_t mesh, tri;
  F3 emission;
  float area;
};
struct SunLight {    // scene.h:43-49: unit direction toward the sun
  F3 dir;
  F3 radiance;       // color * intensity
};
struct Sky {         // environment.h:26-51
  F3 zenith = F3(0.10f, 0.30f, 0.80f);
  F3 horizon = F3(0.55f, 0.70f, 0.85f);
  F3 ground = F3(0.10f, 0.08f, 0.05f);
  float sharpness = 3.0f, intensity = 1.0f;
  F3 eval(F3 d) const {
    if (d.y < 0.f) {
      const float t = stdmin(-d.y, 1.f);
      return intensity * lerp(horizon, ground, t);
    }
    const float t = std::pow(d.y, 1.0f / sharpness);
    return intensity * lerp(horizon, zenith, t);
  }
};
struct Cam {         // camera/camera.h:15-37
  F3 eye, llc, horiz, vert;
  void look_at(F3 e, F3 target, F3 up, float fov, float aspect) {
    const float theta = fov * kPi / 180.f;
    const float h = std::tan(theta / 2.f);
    const float vh = 2.f * h;
    const float vw = aspect * vh;
    const F3 w = normalize(e - target);
    const F3 u = normalize(cross(up, w));
    const F3 v = cross(w, u);
    eye = e;
    horiz = vw * u;
    vert = vh * v;
    llc = eye - horiz * 0.5f - vert * 0.5f - w;
  }
  RayF ray(float s, float t) const {
    RayF r;
    r.o = eye;
    r.d = normalize(llc + s * horiz + t * vert - eye);
    return r;
  }
};

struct Scene {
  std::vector<Mesh> meshes;
  std::vector<Blas> blas;
  std::vector<Inst> pending;   // explicit instances; empty -> one identity instance per mesh
  Tlas tlas;
  std::vector<Mat> mats;
  std::vector<AreaLight> lights;
  std::vector<SunLight> suns;
  Sky sky;
  Cam cam;
  bool built = false;
};

struct PathStats {
  uint64_t closest = 0, shadow = 0, nonfinite = 0;
  // optional log of the path's closest-hit rays (o, d, tmin, tmax) and
  // their hits (t, instance, triangle; -1 ids on a miss), for diagnosis
  std::vector<float>* ray_log = nullptr;
};

// One path for pixel (x, y) and 1-based sample index `spp`
// (cpu_tiled.cpp:20-218).
static F3 radiance(const Scene& sc, int x, int y, int W, int H, int spp, int minb, int maxb, PathStats& st) {
  const int pixel = y * W + x;
  const uint64_t seed = static_cast<uint64_t>(pixel) + static_cast<uint64_t>(spp) * 0xdeadbeefULL;
  Pcg rng = pcg_seeded(seed);
  const float fx = static_cast<float>(x) + rng.next_f32();
  const float fy = static_cast<float>(y) + rng.next_f32();
  RayF ray = sc.cam.ray(fx / static_cast<float>(W), fy / static_cast<float>(H));
  F3 T = F3::splat(1.f);
  F3 L = F3::splat(0.f);
  float prev_pdf = 0.f;
  bool prev_delta = true;
  const int nl = static_cast<int>(sc.lights.size());
  for (int depth = 0; depth < maxb; ++depth) {
    HitF rec;
    st.closest++;
    const bool hit = sc.tlas.closest(ray, rec, sc.blas, nullptr);
    if (st.ray_log) {
      const float e[11] = {ray.o.x, ray.o.y, ray.o.z, ray.d.x, ray.d.y, ray.d.z, ray.tmin, ray.tmax,
                           hit ? rec.t : kInf, hit ? static_cast<float>(rec.inst) : -1.f,
                           hit ? static_cast<float>(rec.prim) : -1.f};
      st.ray_log->insert(st.ray_log->end(), e, e + 11);
    }
    if (!hit) {
      if (sc.sky.intensity > 0.f) L = L + T * sc.sky.eval(ray.d);
      break;
    }
    const Mat& m = sc.mats[rec.mat];
    // emitter hit: first light whose (triangle, mesh) matches (prim, instance)
    bool on_light = false;
    for (const AreaLight& lt : sc.lights) {
      if (lt.tri == static_cast<uint32_t>(rec.prim) && lt.mesh == static_cast<uint32_t>(rec.inst)) {
        float w = 1.0f;
        const float cl = stdmax(0.0f, dot(rec.gn, -ray.d));
        if (cl > 1e-6f) {
          if (depth > 0 && !prev_delta) {
            const float pa = (1.0f / lt.area) * (1.0f / static_cast<float>(nl));
            const float d2 = rec.t * rec.t;
            const float psa = (pa * d2) / cl;
            w = power2(prev_pdf, psa);
          }
          L = L + T * lt.emission * w;
        }
        on_light = true;
        break;
      }
    }
    if (on_light) break;
    // area-light NEE: pick (1 draw), point on triangle (2 draws)
    if (nl > 0 && !m.delta) {
      const float upick = rng.next_f32();
      int li = static_cast<int>(upick * static_cast<float>(nl));
      li = li < 0 ? 0 : (li > nl - 1 ? nl - 1 : li);
      const float pick = 1.f / static_cast<float>(nl);
      const AreaLight& lt = sc.lights[li];
      const Mesh& lm = sc.meshes[lt.mesh];
      F3 p0, p1, p2;
      lm.corners(static_cast<int>(lt.tri), p0, p1, p2);
      const float u1 = rng.next_f32();
      const float u2 = rng.next_f32();
      const float su1 = std::sqrt(u1);
      const float b0 = 1.0f - su1;
      const float b1 = u2 * su1;
      const float b2 = 1.0f - b0 - b1;
      const F3 lp = p0 * b0 + p1 * b1 + p2 * b2;
      const F3 ln = lm.face_normal(static_cast<int>(lt.tri));
      const F3 sd = lp - rec.pos;
      const float d2 = len_sq(sd);
      const float dist = std::sqrt(d2);
      if (dist > 1e-6f) {
        const F3 wi = sd / dist;
        const float cl = stdmax(0.0f, dot(ln, -wi));
        const Frame fr(rec.gn);
        const F3 wo_l = fr.to_local(-ray.d);
        const F3 wi_l = fr.to_local(wi);
        if (cl > 1e-6f && wi_l.z > 1e-6f) {
          const F3 f = nee_f(wo_l, wi_l, m);
          const float bp = nee_pdf(wo_l, wi_l, m);
          const float lpdf = ((1.0f / lt.area) * pick) * (d2 / cl);
          const float w = power2(lpdf, bp);
          const float cs = wi_l.z;
          if (w > 0.0f && lpdf > 1e-12f) {
            RayF sr;
            sr.o = nudge(rec.pos, rec.gn, wi, kShadowOffset);
            sr.d = wi;
            sr.tmin = kShadowOffset;
            sr.tmax = dist - kShadowOffset;
            st.shadow++;
            if (!sc.tlas.any(sr, sc.blas, nullptr)) L = L + T * f * cs * (lt.emission / lpdf) * w;
          }
        }
      }
    }
    // sun NEE with the reference's 1e8 pdf sentinel (cpu_tiled.cpp:148-187)
    if (!sc.suns.empty() && !m.delta) {
      const int ns = static_cast<int>(sc.suns.size());
      const float usun = rng.next_f32();
      int si = static_cast<int>(usun * static_cast<float>(ns));
      si = si < 0 ? 0 : (si > ns - 1 ? ns - 1 : si);
      const float spick = 1.f / static_cast<float>(ns);
      if (spick > 0.f) {
        const SunLight& sun = sc.suns[si];
        const Frame fr(rec.gn);
        const F3 wo_l = fr.to_local(-ray.d);
        const F3 wi_l = fr.to_local(sun.dir);
        if (wi_l.z > 1e-6f) {
          const F3 f = nee_f(wo_l, wi_l, m);
          const float bp = nee_pdf(wo_l, wi_l, m);
          const float cs = wi_l.z;
          const float lpdf = 1e8f * spick;
          const float w = power2(lpdf, bp);
          RayF sr;
          sr.o = nudge(rec.pos, rec.gn, sun.dir, kShadowOffset);
          sr.d = sun.dir;
          sr.tmin = kShadowOffset;
          sr.tmax = kInf;
          st.shadow++;
          if (!sc.tlas.any(sr, sc.blas, nullptr)) L = L + T * f * cs * (sun.radiance / lpdf) * w;
        }
      }
    }
    // Russian roulette
    if (depth >= minb) {
      const float p = stdmax(0.05f, max_comp(T));
      if (rng.next_f32() > p) break;
      T = T * (1.f / p);
    }
    // continuation by BSDF sampling in the shading frame
    const Frame fs(rec.n);
    const F3 wo_l = fs.to_local(-ray.d);
    Sample s;
    if (!sample_bsdf(wo_l, m, rng, s)) break;
    if (s.delta) {
      T = T * s.f;
    } else {
      const float ci = std::fabs(s.wi.z);
      if (s.pdf < 1e-8f || ci < 1e-8f) break;
      T = T * (s.f * ci / s.pdf);
    }
    const F3 wiw = fs.to_world(s.wi);
    ray.o = nudge(rec.pos, rec.gn, wiw, kBounceEps);
    ray.d = wiw;
    ray.tmin = kBounceEps;
    ray.tmax = kInf;
    prev_pdf = s.pdf;
    prev_delta = s.delta;
  }
  if (!std::isfinite(L.x) || !std::isfinite(L.y) || !std::isfinite(L.z)) st.nonfinite++;
  return L;
}

}  // namespace orc

using namespace orc;

static Scene* S(void* h) { return static_cast<Scene*>(h); }

extern "C" {

const char* orc_impl_name(void) { return "oracle-restatement"; }

void* orc_scene_new(void) { return new Scene(); }
void orc_scene_free(void* h) { delete S(h); }

int orc_scene_add_material(void* h, const float* p) {
  Mat m;
  m.base = F3(p[0], p[1], p[2]);
  m.roughness = p[3];
  m.metallic = p[4];
  m.ior = p[5];
  m.transmission = p[6];
  m.subsurface = p[7];
  m.ss_color = F3(p[8], p[9], p[10]);
  m.anisotropy = p[11];
  m.emission = F3(p[12], p[13], p[14]);
  prepare(m);
  S(h)->mats.push_back(m);
  return static_cast<int>(S(h)->mats.size()) - 1;
}

void orc_material_derived(void* h, int i, float* out) {
  const Mat& m = S(h)->mats[i];
  out[0] = m.alpha; out[1] = m.ax; out[2] = m.ay;
  out[3] = m.f0.x; out[4] = m.f0.y; out[5] = m.f0.z;
  out[6] = m.conductor; out[7] = m.transmissive; out[8] = m.sss; out[9] = m.delta; out[10] = m.emissive;
}

int orc_scene_add_mesh(void* h, int nv, const float* pos, const float* nrm, int nt, const int* idx, const int* mat) {
  Mesh m;
  m.p.resize(nv);
  for (int i = 0; i < nv; ++i) m.p[i] = F3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  m.idx.assign(idx, idx + 3 * nt);
  m.mat.assign(mat, mat + nt);
  if (nrm) {
    m.nrm.resize(nv);
    for (int i = 0; i < nv; ++i) m.nrm[i] = F3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
  } else {
    m.smooth_normals();
  }
  S(h)->meshes.push_back(std::move(m));
  return static_cast<int>(S(h)->meshes.size()) - 1;
}

void orc_mesh_normals(void* h, int mesh, float* out) {
  const Mesh& m = S(h)->meshes[mesh];
  for (size_t i = 0; i < m.nrm.size(); ++i) {
    out[3 * i] = m.nrm[i].x; out[3 * i + 1] = m.nrm[i].y; out[3 * i + 2] = m.nrm[i].z;
  }
}

int orc_scene_add_light(void* h, int mesh, int tri, const float* e, float area) {
  S(h)->lights.push_back({static_cast<uint32_t>(mesh), static_cast<uint32_t>(tri), F3(e[0], e[1], e[2]), area});
  return static_cast<int>(S(h)->lights.size()) - 1;
}

// Direction normalised as the scene parser does (scene_file.cpp:137-140).
int orc_scene_add_dir_light(void* h, const float* d, const float* color, float intensity) {
  F3 dir(d[0], d[1], d[2]);
  const float len = length(dir);
  if (len > 1e-6f) dir = dir / len;
  else dir = F3(0.f, 1.f, 0.f);
  S(h)->suns.push_back({dir, F3(color[0], color[1], color[2]) * intensity});
  return static_cast<int>(S(h)->suns.size()) - 1;
}

void orc_scene_set_sky(void* h, const float* zen, const float* hor, const float* gnd, float intensity, float sharp) {
  Sky& s = S(h)->sky;
  s.zenith = F3(zen[0], zen[1], zen[2]);
  s.horizon = F3(hor[0], hor[1], hor[2]);
  s.ground = F3(gnd[0], gnd[1], gnd[2]);
  s.intensity = intensity;
  s.sharpness = sharp;
}

int orc_scene_add_instance(void* h, int mesh, const float* t, float s, const float* rot, int id) {
  Inst in;
  in.mesh = mesh;
  in.id = id;
  in.xf = Affine::trs(F3(t[0], t[1], t[2]), s, F3(rot[0], rot[1], rot[2]));
  S(h)->pending.push_back(in);
  return static_cast<int>(S(h)->pending.size()) - 1;
}

// scene.h:79-121: one BLAS per mesh; without explicit instances every mesh
// gets an identity instance whose id is its mesh index.
int orc_scene_build(void* h) {
  Scene& sc = *S(h);
  sc.blas.clear();
  sc.blas.resize(sc.meshes.size());
  for (size_t i = 0; i < sc.meshes.size(); ++i) sc.blas[i].build(sc.meshes[i]);
  std::vector<Inst> in = sc.pending;
  if (in.empty()) {
    for (size_t i = 0; i < sc.meshes.size(); ++i) {
      Inst x;
      x.mesh = static_cast<int>(i);
      x.id = static_cast<int>(i);
      in.push_back(x);
    }
  }
  sc.tlas.build(std::move(in), sc.blas);
  sc.built = true;
  return 0;
}

void orc_scene_set_camera(void* h, const float* eye, const float* target, const float* up, float fov, float aspect) {
  S(h)->cam.look_at(F3(eye[0], eye[1], eye[2]), F3(target[0], target[1], target[2]), F3(up[0], up[1], up[2]), fov,
                    aspect);
}

void orc_camera_get(void* h, float* out) {
  const Cam& c = S(h)->cam;
  const F3 v[4] = {c.eye, c.llc, c.horiz, c.vert};
  for (int i = 0; i < 4; ++i) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
}

int orc_blas_node_count(void* h, int mesh) { return static_cast<int>(S(h)->blas[mesh].nodes.size()); }
void orc_blas_export(void* h, int mesh, void* nodes, uint32_t* slots) {
  const Blas& b = S(h)->blas[mesh];
  std::memcpy(nodes, b.nodes.data(), b.nodes.size() * sizeof(Node));
  std::memcpy(slots, b.slots.data(), b.slots.size() * sizeof(uint32_t));
}
int orc_tlas_node_count(void* h) { return static_cast<int>(S(h)->tlas.nodes.size()); }
int orc_instance_count(void* h) { return static_cast<int>(S(h)->tlas.inst.size()); }
void orc_tlas_export(void* h, void* nodes, uint32_t* order) {
  const Tlas& t = S(h)->tlas;
  std::memcpy(nodes, t.nodes.data(), t.nodes.size() * sizeof(Node));
  std::memcpy(order, t.order.data(), t.order.size() * sizeof(uint32_t));
}
void orc_instance_export(void* h, float* xf24, int* ids, int* meshes) {
  const Tlas& t = S(h)->tlas;
  for (size_t i = 0; i < t.inst.size(); ++i) {
    std::memcpy(xf24 + 24 * i, t.inst[i].xf.l2w, 12 * sizeof(float));
    std::memcpy(xf24 + 24 * i + 12, t.inst[i].xf.w2l, 12 * sizeof(float));
    ids[i] = t.inst[i].id;
    meshes[i] = t.inst[i].mesh;
  }
}

static RayF load_ray(const float* r) {
  RayF ray;
  ray.o = F3(r[0], r[1], r[2]);
  ray.d = F3(r[3], r[4], r[5]);
  ray.tmin = r[6];
  ray.tmax = r[7];
  return ray;
}

// Closest hit per ray through Scene::intersect (scene.h:126-128): the record
// starts at t = +inf, as a default HitRecord does (ray.h:48).
void orc_trace_closest(void* h, int n, const float* rays, float* hf, int* hi, uint64_t* counters) {
  const Scene& sc = *S(h);
  Counters c;
  for (int i = 0; i < n; ++i) {
    const RayF r = load_ray(rays + 8 * i);
    HitF rec;
    const bool ok = sc.tlas.closest(r, rec, sc.blas, counters ? &c : nullptr);
    float* o = hf + 12 * i;
    int* oi = hi + 4 * i;
    if (!ok) {
      for (int k = 0; k < 12; ++k) o[k] = 0.f;
      o[0] = kInf;
      oi[0] = -1; oi[1] = -1; oi[2] = -1; oi[3] = 0;
      continue;
    }
    const float vals[12] = {rec.t, rec.u, rec.v, rec.pos.x, rec.pos.y, rec.pos.z,
                            rec.n.x, rec.n.y, rec.n.z, rec.gn.x, rec.gn.y, rec.gn.z};
    std::memcpy(o, vals, sizeof(vals));
    oi[0] = rec.inst; oi[1] = rec.prim; oi[2] = rec.mat; oi[3] = rec.front ? 1 : 0;
  }
  if (counters) {
    counters[0] = c.node_tests; counters[1] = c.interior; counters[2] = c.tri_tests; counters[3] = c.inst_visits;
  }
}

void orc_trace_any(void* h, int n, const float* rays, uint8_t* occ, uint64_t* counters) {
  const Scene& sc = *S(h);
  Counters c;
  for (int i = 0; i < n; ++i) occ[i] = sc.tlas.any(load_ray(rays + 8 * i), sc.blas, counters ? &c : nullptr) ? 1 : 0;
  if (counters) {
    counters[0] = c.node_tests; counters[1] = c.interior; counters[2] = c.tri_tests; counters[3] = c.inst_visits;
  }
}

// Adds samples spp_first .. spp_first+spp_count-1 (1-based, the value the
// reference's spp_ holds while rendering that frame) of every pixel into
// accum (W*H*3 floats, pixel order: index y*W+x, y=0 is the bottom row).
// 32x32 tiles over `nthreads` workers, as cpu_tiled.cpp:17 tiles the frame.
// stats: [closest rays, shadow rays, non-finite path radiances].
int orc_render(void* h, int W, int H, int spp_first, int spp_count, int minb, int maxb, int nthreads, float* accum,
               uint64_t* stats) {
  const Scene& sc = *S(h);
  if (!sc.built) return -1;
  if (nthreads < 1) nthreads = 1;
  const int tiles_x = (W + 31) / 32, tiles_y = (H + 31) / 32;
  const int ntiles = tiles_x * tiles_y;
  std::vector<PathStats> per(nthreads);
  for (int spp = spp_first; spp < spp_first + spp_count; ++spp) {
    std::atomic<int> next{0};
    auto work = [&](int tid) {
      PathStats& st = per[tid];
      for (;;) {
        const int tile = next.fetch_add(1);
        if (tile >= ntiles) break;
        const int x0 = (tile % tiles_x) * 32, y0 = (tile / tiles_x) * 32;
        const int x1 = x0 + 32 < W ? x0 + 32 : W, y1 = y0 + 32 < H ? y0 + 32 : H;
        for (int y = y0; y < y1; ++y)
          for (int x = x0; x < x1; ++x) {
            const F3 L = radiance(sc, x, y, W, H, spp, minb, maxb, st);
            float* a = accum + 3 * (static_cast<size_t>(y) * W + x);
            a[0] = a[0] + L.x;
            a[1] = a[1] + L.y;
            a[2] = a[2] + L.z;
          }
      }
    };
    if (nthreads == 1) {
      work(0);
    } else {
      std::vector<std::thread> th;
      for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
      for (auto& t : th) t.join();
    }
  }
  if (stats) {
    PathStats tot;
    for (const auto& p : per) { tot.closest += p.closest; tot.shadow += p.shadow; tot.nonfinite += p.nonfinite; }
    stats[0] = tot.closest; stats[1] = tot.shadow; stats[2] = tot.nonfinite;
  }
  return 0;
}

// Single-path hook for unit tests: radiance of pixel (x, y) at sample spp.
void orc_path(void* h, int W, int H, int x, int y, int spp, int minb, int maxb, float* out3) {
  PathStats st;
  const F3 L = radiance(*S(h), x, y, W, H, spp, minb, maxb, st);
  out3[0] = L.x; out3[1] = L.y; out3[2] = L.z;
}

// The closest-hit rays of one path and their hits, 11 floats each (o, d,
// tmin, tmax, t, instance, triangle); returns the ray count (at most
// max_rays are written).
int orc_path_rays(void* h, int W, int H, int x, int y, int spp, int minb, int maxb, int max_rays, float* out) {
  PathStats st;
  std::vector<float> log;
  st.ray_log = &log;
  radiance(*S(h), x, y, W, H, spp, minb, maxb, st);
  const int n = static_cast<int>(log.size() / 11);
  for (int k = 0; k < n && k < max_rays; ++k)
    for (int j = 0; j < 11; ++j) out[11 * k + j] = log[11 * static_cast<size_t>(k) + j];
  return n;
}

// ---- isolated-function hooks for known-answer tests --------------------------

void orc_pcg_stream(uint64_t seed, uint64_t seq, int n, uint32_t* out) {
  Pcg p = pcg_seeded(seed, seq);
  for (int i = 0; i < n; ++i) out[i] = p.next_u32();
}

// Draws one BSDF sample for a bare material (p: 16 authoring floats) with the
// PCG stream seeded by `seed`; out: wi(3), f(3), pdf, is_delta, ok.
void orc_bsdf_sample(const float* p, const float* wo, uint64_t seed, int n, float* out) {
  Mat m;
  m.base = F3(p[0], p[1], p[2]); m.roughness = p[3]; m.metallic = p[4]; m.ior = p[5];
  m.transmission = p[6]; m.subsurface = p[7]; m.ss_color = F3(p[8], p[9], p[10]);
  m.anisotropy = p[11]; m.emission = F3(p[12], p[13], p[14]);
  prepare(m);
  Pcg rng = pcg_seeded(seed);
  const F3 w(wo[0], wo[1], wo[2]);
  for (int i = 0; i < n; ++i) {
    Sample s;
    const bool ok = sample_bsdf(w, m, rng, s);
    float* o = out + 9 * i;
    o[0] = s.wi.x; o[1] = s.wi.y; o[2] = s.wi.z; o[3] = s.f.x; o[4] = s.f.y; o[5] = s.f.z;
    o[6] = s.pdf; o[7] = s.delta ? 1.f : 0.f; o[8] = ok ? 1.f : 0.f;
  }
}

// NEE eval/pdf for n direction pairs; out: f(3), pdf.
void orc_bsdf_eval(const float* p, int n, const float* wo, const float* wi, float* out) {
  Mat m;
  m.base = F3(p[0], p[1], p[2]); m.roughness = p[3]; m.metallic = p[4]; m.ior = p[5];
  m.transmission = p[6]; m.subsurface = p[7]; m.ss_color = F3(p[8], p[9], p[10]);
  m.anisotropy = p[11]; m.emission = F3(p[12], p[13], p[14]);
  prepare(m);
  for (int i = 0; i < n; ++i) {
    const F3 a(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    const F3 b(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
    const F3 f = nee_f(a, b, m);
    out[4 * i] = f.x; out[4 * i + 1] = f.y; out[4 * i + 2] = f.z;
    out[4 * i + 3] = nee_pdf(a, b, m);
  }
}

}  // extern "C"
