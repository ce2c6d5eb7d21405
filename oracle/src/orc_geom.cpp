// orc_geom.cpp — oracle BVH build and traversal (TEST INFRASTRUCTURE ONLY).
//
// Tree topology depends on the element order produced by libstdc++'s
// std::partition / std::nth_element, which the reference uses
// (blas.cpp:156-169, tlas.cpp:92-104); the same algorithms are called here
// with predicates that return the same truth values, so the trees match
// node for node.
#include "orc_geom.h"

#include <algorithm>
#include <iterator>

namespace orc {

// ---- mesh -----------------------------------------------------------------

// Area-weighted vertex normals, mesh.cpp:12-34: per triangle the raw cross
// product is added to its three corners in order, then each sum normalised.
void Mesh::smooth_normals() {
  nrm.assign(p.size(), F3(0.f, 0.f, 0.f));
  const int nt = tris();
  for (int t = 0; t < nt; ++t) {
    const int a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
    const F3 fn = cross(p[b] - p[a], p[c] - p[a]);
    nrm[a] = nrm[a] + fn;
    nrm[b] = nrm[b] + fn;
    nrm[c] = nrm[c] + fn;
  }
  for (auto& n : nrm) n = normalize(n);
}

// Moller-Trumbore with the reference's parallel cut (|a| < 1e-7) and
// inclusive [tmin, tmax] acceptance (primitives.h:25-49).
static bool mt(const RayF& r, F3 v0, F3 v1, F3 v2, float tmax, float& t, float& u, float& v) {
  const F3 e1 = v1 - v0;
  const F3 e2 = v2 - v0;
  const F3 h = cross(r.d, e2);
  const float a = dot(e1, h);
  if (std::fabs(a) < 1e-7f) return false;
  const float ia = 1.f / a;
  const F3 s = r.o - v0;
  u = ia * dot(s, h);
  if (u < 0.f || u > 1.f) return false;
  const F3 q = cross(s, e1);
  v = ia * dot(r.d, q);
  if (v < 0.f || u + v > 1.f) return false;
  t = ia * dot(e2, q);
  return t >= r.tmin && t <= tmax;
}

// mesh.h:104-129: the hit keeps the raw (not face-forwarded) geometric
// normal and the barycentric blend of vertex normals.
bool Blas::tri_hit(const RayF& r, uint32_t tri, HitF& h) const {
  F3 a, b, c;
  mesh_->corners(static_cast<int>(tri), a, b, c);
  float t, u, v;
  if (!mt(r, a, b, c, h.t, t, u, v)) return false;
  h.t = t;
  h.pos = r.at(t);
  h.u = u;
  h.v = v;
  h.mat = mesh_->mat[tri];
  h.prim = static_cast<int>(tri);
  const F3 gn = normalize(cross(b - a, c - a));
  h.gn = gn;
  h.front = dot(gn, r.d) < 0.f;
  if (!mesh_->nrm.empty()) {
    const int i0 = mesh_->idx[3 * tri], i1 = mesh_->idx[3 * tri + 1], i2 = mesh_->idx[3 * tri + 2];
    h.n = normalize(mesh_->nrm[i0] * (1.f - u - v) + mesh_->nrm[i1] * u + mesh_->nrm[i2] * v);
  } else {
    h.n = gn;
  }
  return true;
}

// ---- BLAS build (binned SAH, 12 bins, leaves of <= 2) ---------------------

void Blas::leaf(uint32_t ni, uint32_t lo, uint32_t hi, const std::vector<Item>& items) {
  nodes[ni].child = lo;
  nodes[ni].count = hi - lo;
  for (uint32_t i = lo; i < hi; ++i) slots[i] = items[i].tri;
}

static void store_box(Node& n, const Box& b) {
  n.lo[0] = b.lo.x; n.lo[1] = b.lo.y; n.lo[2] = b.lo.z;
  n.hi[0] = b.hi.x; n.hi[1] = b.hi.y; n.hi[2] = b.hi.z;
}
static Box load_box(const Node& n) {
  Box b;
  b.lo = F3(n.lo[0], n.lo[1], n.lo[2]);
  b.hi = F3(n.hi[0], n.hi[1], n.hi[2]);
  return b;
}

// Shared binned-SAH split choice for both levels. Returns the split bin, or
// -1 when no finite cost exists. `cost` receives the best normalised cost.
template <class ItemT>
static int sah_bin(const std::vector<ItemT>& items, uint32_t lo, uint32_t hi, int nbins,
                   const Box& cb, int axis, float scale, float parent_area, float& cost) {
  struct Bin { Box b; int n = 0; };
  Bin bins[16];
  for (uint32_t i = lo; i < hi; ++i) {
    int k = static_cast<int>((items[i].c.at(axis) - cb.lo.at(axis)) * scale);
    if (k >= nbins) k = nbins - 1;
    bins[k].n++;
    bins[k].b.grow(items[i].box);
  }
  float la[16], ra[16];
  int ln[16], rn[16];
  Box acc;
  int cnt = 0;
  for (int i = 0; i < nbins - 1; ++i) {
    cnt += bins[i].n;
    ln[i] = cnt;
    acc.grow(bins[i].b);
    la[i] = acc.area();
  }
  Box racc;
  int rcnt = 0;
  for (int i = nbins - 1; i > 0; --i) {
    rcnt += bins[i].n;
    rn[i - 1] = rcnt;
    racc.grow(bins[i].b);
    ra[i - 1] = racc.area();
  }
  cost = kInf;
  int best = -1;
  for (int i = 0; i < nbins - 1; ++i) {
    const float c = (la[i] * static_cast<float>(ln[i]) + ra[i] * static_cast<float>(rn[i])) / parent_area;
    if (c < cost) {
      cost = c;
      best = i;
    }
  }
  return best;
}

void Blas::split(uint32_t ni, uint32_t lo, uint32_t hi, std::vector<Item>& items) {
  Box nb;
  for (uint32_t i = lo; i < hi; ++i) nb.grow(items[i].box);
  store_box(nodes[ni], nb);
  const uint32_t n = hi - lo;
  if (n <= 2) { leaf(ni, lo, hi, items); return; }
  Box cb;
  for (uint32_t i = lo; i < hi; ++i) cb.grow(items[i].c);
  const int axis = max_axis(cb.size());
  const float ext = cb.size().at(axis);
  if (ext == 0.f) { leaf(ni, lo, hi, items); return; }
  const int kBins = 12;
  const float scale = static_cast<float>(kBins) / ext;
  float best_cost;
  const int best = sah_bin(items, lo, hi, kBins, cb, axis, scale, nb.area(), best_cost);
  if (best_cost >= static_cast<float>(n) || best < 0) { leaf(ni, lo, hi, items); return; }
  const float base = cb.lo.at(axis);
  auto it = std::partition(items.begin() + lo, items.begin() + hi, [&](const Item& e) {
    int k = static_cast<int>((e.c.at(axis) - base) * scale);
    if (k >= kBins) k = kBins - 1;
    return k <= best;
  });
  uint32_t mid = static_cast<uint32_t>(std::distance(items.begin(), it));
  if (mid == lo || mid == hi) {
    mid = lo + n / 2;
    std::nth_element(items.begin() + lo, items.begin() + mid, items.begin() + hi,
                     [axis](const Item& a, const Item& b) { return a.c.at(axis) < b.c.at(axis); });
  }
  nodes.emplace_back();
  nodes.emplace_back();
  const uint32_t left = static_cast<uint32_t>(nodes.size() - 2);
  nodes[ni].child = left;
  nodes[ni].count = 0;
  split(left, lo, mid, items);
  split(left + 1, mid, hi, items);
}

void Blas::build(const Mesh& m) {
  mesh_ = &m;
  nodes.clear();
  slots.clear();
  const uint32_t nt = static_cast<uint32_t>(m.tris());
  if (nt == 0) return;
  slots.resize(nt);
  nodes.reserve(2 * static_cast<size_t>(nt));
  std::vector<Item> items(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    items[i].box = m.tri_box(static_cast<int>(i));
    items[i].c = items[i].box.mid();
    items[i].tri = i;
  }
  nodes.emplace_back();
  split(0, 0, nt, items);
}

Box Blas::root_box() const { return nodes.empty() ? Box() : load_box(nodes[0]); }

// ---- BLAS traversal -------------------------------------------------------

// Two-child slab test with SSE MINPS/MAXPS semantics (blas.cpp:25-74):
// near = max(max(max(nx, ny), nz), tmin), far = min(min(min(fx, fy), fz), tmax).
static int pair_test(const Node& a, const Node& b, F3 inv, F3 o, float tmin, float tmax, float tn[2]) {
  int mask = 0;
  const Node* ch[2] = {&a, &b};
  for (int k = 0; k < 2; ++k) {
    const Node& c = *ch[k];
    float nr[3], fr[3];
    const float oo[3] = {o.x, o.y, o.z};
    const float ii[3] = {inv.x, inv.y, inv.z};
    for (int ax = 0; ax < 3; ++ax) {
      const float t0 = (c.lo[ax] - oo[ax]) * ii[ax];
      const float t1 = (c.hi[ax] - oo[ax]) * ii[ax];
      nr[ax] = sse_min(t0, t1);
      fr[ax] = sse_max(t0, t1);
    }
    const float tnear = sse_max(sse_max(sse_max(nr[0], nr[1]), nr[2]), tmin);
    const float tfar = sse_min(sse_min(sse_min(fr[0], fr[1]), fr[2]), tmax);
    tn[k] = tnear;
    if (tnear <= tfar) mask |= (1 << k);
  }
  return mask;
}

// Closest hit, blas.cpp:211-265: the root is never box-tested; interior
// nodes pair-test their children against [tmin, best t]; both hit -> the
// child with the smaller near-t is visited first, ties visit the right child
// first.
bool Blas::closest(const RayF& r, HitF& h, Counters* c) const {
  if (nodes.empty()) return false;
  const F3 inv(1.f / r.d.x, 1.f / r.d.y, 1.f / r.d.z);
  uint32_t stack[64];
  int sp = 0;
  stack[sp++] = 0;
  bool hit = false;
  while (sp > 0) {
    const Node& nd = nodes[stack[--sp]];
    if (nd.count > 0) {
      for (uint32_t i = 0; i < nd.count; ++i) {
        if (c) c->tri_tests++;
        if (tri_hit(r, slots[nd.child + i], h)) hit = true;
      }
      continue;
    }
    const uint32_t l = nd.child, rr = l + 1;
    float tn[2];
    if (c) { c->interior++; c->node_tests += 2; }
    const int m = pair_test(nodes[l], nodes[rr], inv, r.o, r.tmin, h.t, tn);
    if (m == 0) continue;
    if (m == 3) {
      if (tn[0] < tn[1]) { stack[sp++] = rr; stack[sp++] = l; }
      else { stack[sp++] = l; stack[sp++] = rr; }
    } else {
      stack[sp++] = (m & 1) ? l : rr;
    }
  }
  return hit;
}

// Any hit, blas.cpp:269-307.
bool Blas::any(const RayF& r, Counters* c) const {
  if (nodes.empty()) return false;
  const F3 inv(1.f / r.d.x, 1.f / r.d.y, 1.f / r.d.z);
  uint32_t stack[64];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const Node& nd = nodes[stack[--sp]];
    if (nd.count > 0) {
      HitF tmp;
      tmp.t = r.tmax;
      for (uint32_t i = 0; i < nd.count; ++i) {
        if (c) c->tri_tests++;
        if (tri_hit(r, slots[nd.child + i], tmp)) return true;
      }
      continue;
    }
    const uint32_t l = nd.child, rr = l + 1;
    float tn[2];
    if (c) { c->interior++; c->node_tests += 2; }
    const int m = pair_test(nodes[l], nodes[rr], inv, r.o, r.tmin, r.tmax, tn);
    if (m & 2) stack[sp++] = rr;
    if (m & 1) stack[sp++] = l;
  }
  return false;
}

// ---- affine transform (transform.h:24-148) ---------------------------------

Affine Affine::identity() {
  Affine a;
  const float id[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  for (int i = 0; i < 12; ++i) { a.l2w[i] = id[i]; a.w2l[i] = id[i]; }
  return a;
}

// Uniform scale, Euler rotation applied Y then X then Z, then translation.
Affine Affine::trs(F3 t, float s, F3 rot_deg) {
  const F3 rad = rot_deg * (kPi / 180.f);
  const float cx = std::cos(rad.x), sx = std::sin(rad.x);
  const float cy = std::cos(rad.y), sy = std::sin(rad.y);
  const float cz = std::cos(rad.z), sz = std::sin(rad.z);
  const float rxy[3][3] = {{cy, 0.f, sy}, {sx * sy, cx, -sx * cy}, {-cx * sy, sx, cx * cy}};
  const float rz[3][3] = {{cz, -sz, 0.f}, {sz, cz, 0.f}, {0.f, 0.f, 1.f}};
  float m[3][3] = {};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) m[r][c] += rz[r][k] * rxy[k][c];
  Affine a;
  for (int j = 0; j < 3; ++j) {
    a.l2w[3 * j + 0] = m[0][j] * s;
    a.l2w[3 * j + 1] = m[1][j] * s;
    a.l2w[3 * j + 2] = m[2][j] * s;
  }
  a.l2w[9] = t.x; a.l2w[10] = t.y; a.l2w[11] = t.z;
  const float is = 1.f / s;
  for (int j = 0; j < 3; ++j) {
    a.w2l[3 * j + 0] = m[j][0] * is;
    a.w2l[3 * j + 1] = m[j][1] * is;
    a.w2l[3 * j + 2] = m[j][2] * is;
  }
  for (int i = 0; i < 3; ++i)
    a.w2l[9 + i] = -(m[0][i] * t.x + m[1][i] * t.y + m[2][i] * t.z) * is;
  return a;
}

static F3 col(const float* m, int j) { return F3(m[3 * j], m[3 * j + 1], m[3 * j + 2]); }
static F3 xf_point(const float* m, F3 p) {
  return col(m, 0) * p.x + col(m, 1) * p.y + col(m, 2) * p.z + col(m, 3);
}
static F3 xf_dir(const float* m, F3 d) { return col(m, 0) * d.x + col(m, 1) * d.y + col(m, 2) * d.z; }

// The local direction is left unnormalised so t is shared by both spaces.
RayF ray_to_local(const Affine& a, const RayF& w) {
  RayF l;
  l.o = xf_point(a.w2l, w.o);
  l.d = xf_dir(a.w2l, w.d);
  l.tmin = w.tmin;
  l.tmax = w.tmax;
  return l;
}

// (A^-T) n through the rows of world_to_local, then renormalised.
F3 normal_to_world(const Affine& a, F3 n) {
  const F3 c0 = col(a.w2l, 0), c1 = col(a.w2l, 1), c2 = col(a.w2l, 2);
  return normalize(F3(c0.x * n.x + c0.y * n.y + c0.z * n.z, c1.x * n.x + c1.y * n.y + c1.z * n.z,
                      c2.x * n.x + c2.y * n.y + c2.z * n.z));
}

// ---- TLAS (binned SAH, 8 bins, one instance per leaf) ----------------------

void Tlas::leaf(uint32_t ni, uint32_t lo, uint32_t hi, const std::vector<Item>& items) {
  nodes[ni].child = lo;
  nodes[ni].count = hi - lo;
  for (uint32_t i = lo; i < hi; ++i) order[i] = items[i].ii;
}

void Tlas::split(uint32_t ni, uint32_t lo, uint32_t hi, std::vector<Item>& items) {
  Box nb;
  for (uint32_t i = lo; i < hi; ++i) nb.grow(items[i].box);
  store_box(nodes[ni], nb);
  const uint32_t n = hi - lo;
  if (n <= 1) { leaf(ni, lo, hi, items); return; }
  Box cb;
  for (uint32_t i = lo; i < hi; ++i) cb.grow(items[i].c);
  const int axis = max_axis(cb.size());
  const float ext = cb.size().at(axis);
  if (ext == 0.f) { leaf(ni, lo, hi, items); return; }
  const int kBins = 8;
  const float scale = static_cast<float>(kBins) / ext;
  float best_cost;
  const int best = sah_bin(items, lo, hi, kBins, cb, axis, scale, nb.area(), best_cost);
  if (best < 0) { leaf(ni, lo, hi, items); return; }
  const float base = cb.lo.at(axis);
  auto it = std::partition(items.begin() + lo, items.begin() + hi, [&](const Item& e) {
    int k = static_cast<int>((e.c.at(axis) - base) * scale);
    if (k >= kBins) k = kBins - 1;
    return k <= best;
  });
  uint32_t mid = static_cast<uint32_t>(std::distance(items.begin(), it));
  if (mid == lo || mid == hi) {
    mid = lo + n / 2;
    std::nth_element(items.begin() + lo, items.begin() + mid, items.begin() + hi,
                     [axis](const Item& a, const Item& b) { return a.c.at(axis) < b.c.at(axis); });
  }
  nodes.emplace_back();
  nodes.emplace_back();
  const uint32_t left = static_cast<uint32_t>(nodes.size() - 2);
  nodes[ni].child = left;
  nodes[ni].count = 0;
  split(left, lo, mid, items);
  split(left + 1, mid, hi, items);
}

static Box world_box(const Affine& a, const Box& b) {
  Box wb;
  for (int i = 0; i < 8; ++i) {
    const F3 p((i & 1) ? b.hi.x : b.lo.x, (i & 2) ? b.hi.y : b.lo.y, (i & 4) ? b.hi.z : b.lo.z);
    wb.grow(xf_point(a.l2w, p));
  }
  return wb;
}

void Tlas::build(std::vector<Inst> in, const std::vector<Blas>& blas) {
  inst = std::move(in);
  nodes.clear();
  order.clear();
  const uint32_t n = static_cast<uint32_t>(inst.size());
  if (n == 0) return;
  order.resize(n);
  std::vector<Item> items(n);
  for (uint32_t i = 0; i < n; ++i) {
    inst[i].world = world_box(inst[i].xf, blas[inst[i].mesh].root_box());
    items[i].box = inst[i].world;
    items[i].c = inst[i].world.mid();
    items[i].ii = i;
  }
  nodes.emplace_back();
  split(0, 0, n, items);
}

// Sign-selected slab test with std::max/std::min initializer-list folds
// (aabb.h:59-71).
static bool slab_fast(const Node& nd, F3 inv, F3 o, const int sg[3], float tmin, float tmax) {
  const float* lohi[2] = {nd.lo, nd.hi};
  const float txn = (lohi[sg[0]][0] - o.x) * inv.x;
  const float txf = (lohi[1 - sg[0]][0] - o.x) * inv.x;
  const float tyn = (lohi[sg[1]][1] - o.y) * inv.y;
  const float tyf = (lohi[1 - sg[1]][1] - o.y) * inv.y;
  const float tzn = (lohi[sg[2]][2] - o.z) * inv.z;
  const float tzf = (lohi[1 - sg[2]][2] - o.z) * inv.z;
  const float lo = max_first(max_first(max_first(tmin, txn), tyn), tzn);
  const float hi = min_first(min_first(min_first(tmax, txf), tyf), tzf);
  return lo <= hi;
}

// Closest hit, tlas.cpp:172-245: test-on-pop against [tmin, best t]; leaf
// instances transform the ray and descend their BLAS with the running best
// record; interior children are ordered by their box centres projected on
// the parent's longest axis.
bool Tlas::closest(const RayF& r, HitF& h, const std::vector<Blas>& blas, Counters* c) const {
  if (nodes.empty()) return false;
  const F3 inv(1.f / r.d.x, 1.f / r.d.y, 1.f / r.d.z);
  const int sg[3] = {r.d.x < 0, r.d.y < 0, r.d.z < 0};
  uint32_t stack[64];
  int sp = 0;
  stack[sp++] = 0;
  bool hit = false;
  while (sp > 0) {
    const Node& nd = nodes[stack[--sp]];
    if (c) c->node_tests++;
    if (!slab_fast(nd, inv, r.o, sg, r.tmin, h.t)) continue;
    if (nd.count > 0) {
      for (uint32_t i = 0; i < nd.count; ++i) {
        const Inst& in = inst[order[nd.child + i]];
        if (c) c->inst_visits++;
        const RayF lr = ray_to_local(in.xf, r);
        HitF lh = h;
        if (blas[in.mesh].closest(lr, lh, c)) {
          lh.pos = r.at(lh.t);
          lh.n = normal_to_world(in.xf, lh.n);
          lh.gn = normal_to_world(in.xf, lh.gn);
          lh.front = dot(lh.gn, r.d) < 0.f;
          lh.inst = in.id;
          h = lh;
          hit = true;
        }
      }
    } else {
      if (c) c->interior++;
      const uint32_t l = nd.child, rr = l + 1;
      const int ax = max_axis(load_box(nd).size());
      const float oo = r.o.at(ax), ii = inv.at(ax);
      const float tl = (load_box(nodes[l]).mid().at(ax) - oo) * ii;
      const float tr = (load_box(nodes[rr]).mid().at(ax) - oo) * ii;
      if (tl < tr) { stack[sp++] = rr; stack[sp++] = l; }
      else { stack[sp++] = l; stack[sp++] = rr; }
    }
  }
  return hit;
}

// Any hit, tlas.cpp:248-290.
bool Tlas::any(const RayF& r, const std::vector<Blas>& blas, Counters* c) const {
  if (nodes.empty()) return false;
  const F3 inv(1.f / r.d.x, 1.f / r.d.y, 1.f / r.d.z);
  const int sg[3] = {r.d.x < 0, r.d.y < 0, r.d.z < 0};
  uint32_t stack[64];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const Node& nd = nodes[stack[--sp]];
    if (c) c->node_tests++;
    if (!slab_fast(nd, inv, r.o, sg, r.tmin, r.tmax)) continue;
    if (nd.count > 0) {
      for (uint32_t i = 0; i < nd.count; ++i) {
        const Inst& in = inst[order[nd.child + i]];
        if (c) c->inst_visits++;
        if (blas[in.mesh].any(ray_to_local(in.xf, r), c)) return true;
      }
    } else {
      if (c) c->interior++;
      stack[sp++] = nd.child;
      stack[sp++] = nd.child + 1;
    }
  }
  return false;
}

}  // namespace orc
