// scene.cpp — host scene builder of the HIP backend (compiled with g++,
// -ffp-contract=off, so every float operation rounds as the reference's).
//
// Trees are built with an explicit LIFO task stack that visits nodes in the
// same pre-order as the reference's recursive subdivide (children are
// allocated as an adjacent pair when their parent is split, the left
// subtree completes before the right one starts), and with libstdc++'s
// std::partition / std::nth_element, whose element order the reference's
// leaf order inherits (geometry/blas.cpp:156-169, geometry/tlas.cpp:92-104).
#include "scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace xnh {

namespace {
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float comp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 unit(V3 v) {
  const float len = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  const float r = 1.f / len;
  return {v.x * r, v.y * r, v.z * r};
}
inline int longest_axis(V3 e) { return (e.x > e.y && e.x > e.z) ? 0 : (e.y > e.z ? 1 : 2); }
inline float keep_min(float cur, float v) { return (v < cur) ? v : cur; }
inline float keep_max(float cur, float v) { return (cur < v) ? v : cur; }
constexpr float kPiF = 3.14159265358979323846f;
}  // namespace

float Aabb::kBig() { return std::numeric_limits<float>::infinity(); }
void Aabb::add(const V3& p) {
  lo = {keep_min(lo.x, p.x), keep_min(lo.y, p.y), keep_min(lo.z, p.z)};
  hi = {keep_max(hi.x, p.x), keep_max(hi.y, p.y), keep_max(hi.z, p.z)};
}
void Aabb::add(const Aabb& b) {
  lo = {keep_min(lo.x, b.lo.x), keep_min(lo.y, b.lo.y), keep_min(lo.z, b.lo.z)};
  hi = {keep_max(hi.x, b.hi.x), keep_max(hi.y, b.hi.y), keep_max(hi.z, b.hi.z)};
}
V3 Aabb::center() const { return (lo + hi) * 0.5f; }
V3 Aabb::extent() const { return hi - lo; }
float Aabb::half_area2() const {
  const V3 e = extent();
  return 2.f * (e.x * e.y + e.y * e.z + e.z * e.x);
}

HostScene::HostScene() {
  const float z[3] = {0.10f, 0.30f, 0.80f}, h[3] = {0.55f, 0.70f, 0.85f}, g[3] = {0.10f, 0.08f, 0.05f};
  std::memcpy(sky.zenith, z, sizeof z);
  std::memcpy(sky.horizon, h, sizeof h);
  std::memcpy(sky.ground, g, sizeof g);
  sky.sharpness = 3.0f;
  sky.intensity = 1.0f;
}

// material/material.h:55-75
xnh_material prepare_material(const float* p) {
  xnh_material m{};
  const float rough = p[3], metallic = p[4], ior = p[5], transmission = p[6], sss = p[7], aniso = p[11];
  m.base_color[0] = p[0]; m.base_color[1] = p[1]; m.base_color[2] = p[2];
  m.metallic = metallic;
  m.ior = ior;
  m.alpha = keep_max(rough * rough, 1e-6f);
  const float aspect = std::sqrt(1.0f - 0.9f * std::fabs(aniso));
  m.alpha_x = keep_max(m.alpha / aspect, 1e-6f);
  m.alpha_y = keep_max(m.alpha * aspect, 1e-6f);
  const float fd = ((ior - 1.0f) * (ior - 1.0f)) / ((ior + 1.0f) * (ior + 1.0f));
  for (int c = 0; c < 3; ++c) m.f0[c] = fd + metallic * (p[c] - fd);
  m.subsurface = sss;
  m.subsurface_color[0] = p[8]; m.subsurface_color[1] = p[9]; m.subsurface_color[2] = p[10];
  uint32_t f = 0;
  if (metallic > 0.5f) f |= XNH_MAT_CONDUCTOR;
  if (transmission > 0.01f) f |= XNH_MAT_TRANSMISSIVE;
  if (sss > 0.01f) f |= XNH_MAT_SUBSURFACE;
  if (m.alpha < 0.001f) f |= XNH_MAT_DELTA;
  if (p[12] > 0.0f || p[13] > 0.0f || p[14] > 0.0f) f |= XNH_MAT_EMISSIVE;
  m.flags = f;
  return m;
}

// geometry/mesh.cpp:12-34: raw face cross products summed into the corners
// in triangle order, then normalised.
void smooth_normals(HostMesh& m) {
  const size_t nv = m.vx.size();
  m.nx.assign(nv, 0.f);
  m.ny.assign(nv, 0.f);
  m.nz.assign(nv, 0.f);
  const int nt = m.tris();
  for (int t = 0; t < nt; ++t) {
    const int i[3] = {m.indices[3 * t], m.indices[3 * t + 1], m.indices[3 * t + 2]};
    const V3 p0{m.vx[i[0]], m.vy[i[0]], m.vz[i[0]]};
    const V3 p1{m.vx[i[1]], m.vy[i[1]], m.vz[i[1]]};
    const V3 p2{m.vx[i[2]], m.vy[i[2]], m.vz[i[2]]};
    const V3 fn = cross(p1 - p0, p2 - p0);
    for (int k = 0; k < 3; ++k) {
      m.nx[i[k]] += fn.x;
      m.ny[i[k]] += fn.y;
      m.nz[i[k]] += fn.z;
    }
  }
  for (size_t v = 0; v < nv; ++v) {
    const V3 n = unit({m.nx[v], m.ny[v], m.nz[v]});
    m.nx[v] = n.x;
    m.ny[v] = n.y;
    m.nz[v] = n.z;
  }
}

namespace {

struct Item {
  Aabb box;
  V3 c;
  uint32_t id;
};

struct Task {
  uint32_t node, lo, hi;
};

// Binned SAH over [lo, hi) (blas.cpp:97-180 with 12 bins and leaves of at
// most 2; tlas.cpp:40-113 with 8 bins, leaves of 1 and no leaf-cost test).
template <int kBins, uint32_t kLeafMax, bool kLeafCost>
void build_tree(std::vector<Item>& items, std::vector<xnh_bvh_node>& nodes, std::vector<uint32_t>& slot_ids) {
  const uint32_t n_items = static_cast<uint32_t>(items.size());
  nodes.clear();
  slot_ids.assign(n_items, 0u);
  if (n_items == 0) return;
  nodes.reserve(2 * static_cast<size_t>(n_items));
  nodes.push_back(xnh_bvh_node{});
  std::vector<Task> todo;
  todo.push_back({0u, 0u, n_items});
  auto make_leaf = [&](const Task& t) {
    nodes[t.node].child = t.lo;
    nodes[t.node].count = t.hi - t.lo;
    for (uint32_t i = t.lo; i < t.hi; ++i) slot_ids[i] = items[i].id;
  };
  while (!todo.empty()) {
    const Task t = todo.back();
    todo.pop_back();
    Aabb box;
    for (uint32_t i = t.lo; i < t.hi; ++i) box.add(items[i].box);
    xnh_bvh_node& nd = nodes[t.node];
    nd.lo[0] = box.lo.x; nd.lo[1] = box.lo.y; nd.lo[2] = box.lo.z;
    nd.hi[0] = box.hi.x; nd.hi[1] = box.hi.y; nd.hi[2] = box.hi.z;
    const uint32_t count = t.hi - t.lo;
    if (count <= kLeafMax) { make_leaf(t); continue; }
    Aabb cb;
    for (uint32_t i = t.lo; i < t.hi; ++i) cb.add(items[i].c);
    const int axis = longest_axis(cb.extent());
    const float span = comp(cb.extent(), axis);
    if (span == 0.f) { make_leaf(t); continue; }
    const float origin = comp(cb.lo, axis);
    const float scale = static_cast<float>(kBins) / span;
    auto bin_of = [&](const Item& it) {
      int b = static_cast<int>((comp(it.c, axis) - origin) * scale);
      return b >= kBins ? kBins - 1 : b;
    };
    Aabb bin_box[kBins];
    int bin_n[kBins] = {};
    for (uint32_t i = t.lo; i < t.hi; ++i) {
      const int b = bin_of(items[i]);
      bin_n[b]++;
      bin_box[b].add(items[i].box);
    }
    float area_l[kBins - 1], area_r[kBins - 1];
    int n_l[kBins - 1], n_r[kBins - 1];
    {
      Aabb acc;
      int run = 0;
      for (int i = 0; i < kBins - 1; ++i) {
        run += bin_n[i];
        n_l[i] = run;
        acc.add(bin_box[i]);
        area_l[i] = acc.half_area2();
      }
      Aabb racc;
      int rrun = 0;
      for (int i = kBins - 1; i > 0; --i) {
        rrun += bin_n[i];
        n_r[i - 1] = rrun;
        racc.add(bin_box[i]);
        area_r[i - 1] = racc.half_area2();
      }
    }
    const float parent = box.half_area2();
    float best_cost = std::numeric_limits<float>::infinity();
    int best = -1;
    for (int i = 0; i < kBins - 1; ++i) {
      const float c = (area_l[i] * static_cast<float>(n_l[i]) + area_r[i] * static_cast<float>(n_r[i])) / parent;
      if (c < best_cost) {
        best_cost = c;
        best = i;
      }
    }
    if (best < 0 || (kLeafCost && best_cost >= static_cast<float>(count))) { make_leaf(t); continue; }
    auto first = items.begin() + t.lo, last = items.begin() + t.hi;
    uint32_t mid = static_cast<uint32_t>(
        std::partition(first, last, [&](const Item& it) { return bin_of(it) <= best; }) - items.begin());
    if (mid == t.lo || mid == t.hi) {
      mid = t.lo + count / 2;
      std::nth_element(first, items.begin() + mid, last,
                       [axis](const Item& a, const Item& b) { return comp(a.c, axis) < comp(b.c, axis); });
    }
    const uint32_t left = static_cast<uint32_t>(nodes.size());
    nodes.push_back(xnh_bvh_node{});
    nodes.push_back(xnh_bvh_node{});
    nodes[t.node].child = left;
    nodes[t.node].count = 0;
    todo.push_back({left + 1, mid, t.hi});
    todo.push_back({left, t.lo, mid});
  }
}

Aabb tri_box(const HostMesh& m, int t) {
  Aabb b;
  for (int k = 0; k < 3; ++k) {
    const int v = m.indices[3 * t + k];
    b.add(V3{m.vx[v], m.vy[v], m.vz[v]});
  }
  return b;
}

V3 xf_point(const float* m, V3 p) {
  const V3 c0{m[0], m[1], m[2]}, c1{m[3], m[4], m[5]}, c2{m[6], m[7], m[8]}, c3{m[9], m[10], m[11]};
  return c0 * p.x + c1 * p.y + c2 * p.z + c3;
}

}  // namespace

void build_blas(HostMesh& m) {
  const int nt = m.tris();
  std::vector<Item> items(nt);
  for (int t = 0; t < nt; ++t) {
    items[t].box = tri_box(m, t);
    items[t].c = items[t].box.center();
    items[t].id = static_cast<uint32_t>(t);
  }
  build_tree<12, 2, true>(items, m.nodes, m.prims);
}

namespace {

// AABB of an analytic primitive, as the primitives' aabb() compute it
// (geometry/primitives.h:96-99, 129-145, 204-206, 254-265, 308-317).
Aabb prim_box(const xnh_prim& pr) {
  const float* q = pr.p;
  Aabb b;
  switch (pr.type) {
    case XNH_PRIM_SPHERE: {
      const V3 c{q[0], q[1], q[2]}, r{q[3], q[3], q[3]};
      b.lo = c - r;
      b.hi = c + r;
      break;
    }
    case XNH_PRIM_PLANE: {
      const float kHuge = 1e4f, off = q[3];
      const V3 n{q[0], q[1], q[2]};
      const V3 a = n * (off - 0.01f), c = n * (off + 0.01f);
      float lo[3] = {-kHuge + a.x, -kHuge + a.y, -kHuge + a.z};
      float hi[3] = {kHuge + c.x, kHuge + c.y, kHuge + c.z};
      for (int i = 0; i < 3; ++i) {
        if (std::abs(q[i]) > 0.9f) {
          lo[i] = off - 0.01f;
          hi[i] = off + 0.01f;
        }
      }
      b.lo = {lo[0], lo[1], lo[2]};
      b.hi = {hi[0], hi[1], hi[2]};
      break;
    }
    case XNH_PRIM_BOX: {
      const V3 c{q[0], q[1], q[2]}, h{q[3], q[4], q[5]};
      b.lo = c - h;
      b.hi = c + h;
      break;
    }
    case XNH_PRIM_DISK: {
      const V3 c{q[0], q[1], q[2]};
      const float r = q[6];
      float e[3] = {r * std::sqrt(1.f - q[3] * q[3]), r * std::sqrt(1.f - q[4] * q[4]),
                    r * std::sqrt(1.f - q[5] * q[5])};
      for (float& v : e) v = keep_max(v, 0.001f);   // max3(e, Vec3(0.001f))
      b.lo = c - V3{e[0], e[1], e[2]};
      b.hi = c + V3{e[0], e[1], e[2]};
      break;
    }
    default: {   // XNH_PRIM_QUAD
      const V3 o{q[0], q[1], q[2]}, u{q[3], q[4], q[5]}, v{q[6], q[7], q[8]};
      b.add(o);
      b.add(o + u);
      b.add(o + v);
      b.add(o + u + v);
      b.lo = b.lo - V3{0.001f, 0.001f, 0.001f};
      b.hi = b.hi + V3{0.001f, 0.001f, 0.001f};
      break;
    }
  }
  return b;
}

}  // namespace

// PrimBLAS::build (geometry/prim_blas.cpp:142-160, 34-118): the binned SAH
// of the triangle BLAS over the primitives' boxes.
void build_prims(HostScene& s) {
  const uint32_t n = static_cast<uint32_t>(s.prims.size());
  std::vector<Item> items(n);
  for (uint32_t i = 0; i < n; ++i) {
    items[i].box = prim_box(s.prims[i]);
    items[i].c = items[i].box.center();
    items[i].id = i;
  }
  build_tree<12, 2, true>(items, s.prim_nodes, s.prim_order);
}

// geometry/transform.h:105-148: uniform scale, Euler Y then X then Z.
xnh_instance make_instance(int32_t mesh, int32_t id, const float t[3], float s, const float rot_deg[3]) {
  xnh_instance in{};
  in.mesh = mesh;
  in.instance_id = id;
  const float k = kPiF / 180.f;
  const float ax = rot_deg[0] * k, ay = rot_deg[1] * k, az = rot_deg[2] * k;
  const float cx = std::cos(ax), sx = std::sin(ax), cy = std::cos(ay), sy = std::sin(ay);
  const float cz = std::cos(az), sz = std::sin(az);
  const float a[3][3] = {{cy, 0.f, sy}, {sx * sy, cx, -sx * cy}, {-cx * sy, sx, cx * cy}};
  const float b[3][3] = {{cz, -sz, 0.f}, {sz, cz, 0.f}, {0.f, 0.f, 1.f}};
  float m[3][3] = {};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      for (int q = 0; q < 3; ++q) m[r][c] += b[r][q] * a[q][c];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) in.local_to_world[3 * j + i] = m[i][j] * s;
  in.local_to_world[9] = t[0];
  in.local_to_world[10] = t[1];
  in.local_to_world[11] = t[2];
  const float is = 1.f / s;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) in.world_to_local[3 * j + i] = m[j][i] * is;
  for (int i = 0; i < 3; ++i) in.world_to_local[9 + i] = -(m[0][i] * t[0] + m[1][i] * t[1] + m[2][i] * t[2]) * is;
  return in;
}

void build_tlas(HostScene& s) {
  s.instances.clear();
  if (s.pending.empty()) {
    const float zero[3] = {0.f, 0.f, 0.f};
    for (size_t i = 0; i < s.meshes.size(); ++i) {
      xnh_instance in{};
      in.mesh = static_cast<int32_t>(i);
      in.instance_id = static_cast<int32_t>(i);
      const float id[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
      std::memcpy(in.local_to_world, id, sizeof id);
      std::memcpy(in.world_to_local, id, sizeof id);
      (void)zero;
      s.instances.push_back(in);
    }
  } else {
    for (const auto& p : s.pending) s.instances.push_back(make_instance(p.mesh, p.id, p.t, p.s, p.rot));
  }
  // the primitives: one identity instance after the others
  // (Scene::build_acceleration, scene/scene.h:97-110)
  if (!s.prims.empty()) {
    xnh_instance in{};
    in.mesh = XNH_PRIM_GEOMETRY;
    in.instance_id = static_cast<int32_t>(s.instances.size());
    const float id[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    std::memcpy(in.local_to_world, id, sizeof id);
    std::memcpy(in.world_to_local, id, sizeof id);
    s.instances.push_back(in);
  }
  const uint32_t n = static_cast<uint32_t>(s.instances.size());
  std::vector<Item> items(n);
  for (uint32_t i = 0; i < n; ++i) {
    const std::vector<xnh_bvh_node>& nodes =
        s.instances[i].mesh == XNH_PRIM_GEOMETRY ? s.prim_nodes : s.meshes[s.instances[i].mesh].nodes;
    Aabb local;
    if (!nodes.empty()) {
      local.lo = {nodes[0].lo[0], nodes[0].lo[1], nodes[0].lo[2]};
      local.hi = {nodes[0].hi[0], nodes[0].hi[1], nodes[0].hi[2]};
    }
    Aabb w;   // the eight transformed corners (transform.h:88-98)
    for (int c = 0; c < 8; ++c) {
      const V3 p{(c & 1) ? local.hi.x : local.lo.x, (c & 2) ? local.hi.y : local.lo.y,
                 (c & 4) ? local.hi.z : local.lo.z};
      w.add(xf_point(s.instances[i].local_to_world, p));
    }
    items[i].box = w;
    items[i].c = w.center();
    items[i].id = i;
  }
  build_tree<8, 1, false>(items, s.tlas_nodes, s.tlas_order);
}

// camera/camera.h:15-32
xnh_camera look_at(const float eye[3], const float target[3], const float up[3], float fov_deg, float aspect) {
  const float theta = fov_deg * kPiF / 180.f;
  const float h = std::tan(theta / 2.f);
  const float vh = 2.f * h;
  const float vw = aspect * vh;
  const V3 e{eye[0], eye[1], eye[2]}, tg{target[0], target[1], target[2]}, u0{up[0], up[1], up[2]};
  const V3 w = unit(e - tg);
  const V3 u = unit(cross(u0, w));
  const V3 v = cross(w, u);
  const V3 hz = u * vw, vt = v * vh;
  const V3 llc = e - hz * 0.5f - vt * 0.5f - w;
  xnh_camera c{};
  const V3 all[4] = {e, llc, hz, vt};
  float* dst[4] = {c.eye, c.lower_left, c.horizontal, c.vertical};
  for (int i = 0; i < 4; ++i) { dst[i][0] = all[i].x; dst[i][1] = all[i].y; dst[i][2] = all[i].z; }
  return c;
}

}  // namespace xnh
