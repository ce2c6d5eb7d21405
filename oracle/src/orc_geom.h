// orc_geom.h — meshes, binned-SAH BVHs, instancing for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
//
// Restates, in the reference's float evaluation order:
//   Moller-Trumbore           geometry/primitives.h:25-49
//   mesh hit record           geometry/mesh.h:104-129
//   smooth normals            geometry/mesh.cpp:12-34
//   BLAS build / traversal    geometry/blas.cpp:25-307
//   TLAS build / traversal    geometry/tlas.cpp:21-290
//   affine TRS transform      geometry/transform.h:24-148
//   slab test                 geometry/aabb.h:59-71
#pragma once
#include <cstdint>
#include <vector>

#include "orc_math.h"

namespace orc {

struct Box {
  F3 lo = F3::splat(kInf), hi = F3::splat(-kInf);
  void grow(F3 p) {  // aabb.h:22-25 via min3/max3 (std::min/std::max)
    lo = F3(p.x < lo.x ? p.x : lo.x, p.y < lo.y ? p.y : lo.y, p.z < lo.z ? p.z : lo.z);
    hi = F3(hi.x < p.x ? p.x : hi.x, hi.y < p.y ? p.y : hi.y, hi.z < p.z ? p.z : hi.z);
  }
  void grow(const Box& b) {
    lo = F3(b.lo.x < lo.x ? b.lo.x : lo.x, b.lo.y < lo.y ? b.lo.y : lo.y, b.lo.z < lo.z ? b.lo.z : lo.z);
    hi = F3(hi.x < b.hi.x ? b.hi.x : hi.x, hi.y < b.hi.y ? b.hi.y : hi.y, hi.z < b.hi.z ? b.hi.z : hi.z);
  }
  F3 mid() const { return (lo + hi) * 0.5f; }
  F3 size() const { return hi - lo; }
  float area() const {
    const F3 e = size();
    return 2.f * (e.x * e.y + e.y * e.z + e.z * e.x);
  }
};

// 32-byte node, binary-identical to the reference's BLASNode / TLASNode
// (blas.h:21-28, tlas.h:91-96): box, child (left child, right = left+1,
// or first slot of a leaf) and count (0 = interior).
struct Node {
  float lo[3], hi[3];
  uint32_t child, count;
};
static_assert(sizeof(Node) == 32, "node layout");

struct Mesh {
  std::vector<F3> p;       // vertex positions
  std::vector<F3> nrm;     // vertex normals (always present after build)
  std::vector<int32_t> idx;
  std::vector<int32_t> mat;
  int tris() const { return static_cast<int>(idx.size() / 3); }
  void corners(int t, F3& a, F3& b, F3& c) const {
    a = p[idx[3 * t]];
    b = p[idx[3 * t + 1]];
    c = p[idx[3 * t + 2]];
  }
  F3 face_normal(int t) const {
    F3 a, b, c;
    corners(t, a, b, c);
    return normalize(cross(b - a, c - a));
  }
  Box tri_box(int t) const {
    F3 a, b, c;
    corners(t, a, b, c);
    Box bx;
    bx.grow(a);
    bx.grow(b);
    bx.grow(c);
    return bx;
  }
  void smooth_normals();
};

struct Counters {
  uint64_t node_tests = 0;   // boxes tested (BLAS pair test counts 2, TLAS pop counts 1)
  uint64_t interior = 0;     // interior nodes expanded
  uint64_t tri_tests = 0;    // Moller-Trumbore evaluations
  uint64_t inst_visits = 0;  // instance ray transforms
};

class Blas {
 public:
  void build(const Mesh& m);
  bool closest(const RayF& r, HitF& h, Counters* c) const;
  bool any(const RayF& r, Counters* c) const;
  Box root_box() const;
  std::vector<Node> nodes;
  std::vector<uint32_t> slots;  // leaf slot -> mesh triangle
 private:
  struct Item { Box box; F3 c; uint32_t tri; };
  void split(uint32_t ni, uint32_t lo, uint32_t hi, std::vector<Item>& items);
  void leaf(uint32_t ni, uint32_t lo, uint32_t hi, const std::vector<Item>& items);
  bool tri_hit(const RayF& r, uint32_t tri, HitF& h) const;
  const Mesh* mesh_ = nullptr;
};

struct Affine {
  float l2w[12];  // column-major 3x4: cols 0..2 basis, col 3 translation
  float w2l[12];
  static Affine identity();
  static Affine trs(F3 t, float s, F3 rot_deg);
};

struct Inst {
  int mesh = 0;
  int id = 0;
  Affine xf = Affine::identity();
  Box world;
};

class Tlas {
 public:
  void build(std::vector<Inst> in, const std::vector<Blas>& blas);
  bool closest(const RayF& r, HitF& h, const std::vector<Blas>& blas, Counters* c) const;
  bool any(const RayF& r, const std::vector<Blas>& blas, Counters* c) const;
  std::vector<Node> nodes;
  std::vector<uint32_t> order;  // leaf slot -> instance index
  std::vector<Inst> inst;
 private:
  struct Item { Box box; F3 c; uint32_t ii; };
  void split(uint32_t ni, uint32_t lo, uint32_t hi, std::vector<Item>& items);
  void leaf(uint32_t ni, uint32_t lo, uint32_t hi, const std::vector<Item>& items);
};

RayF ray_to_local(const Affine& a, const RayF& w);
F3 normal_to_world(const Affine& a, F3 n);

}  // namespace orc
