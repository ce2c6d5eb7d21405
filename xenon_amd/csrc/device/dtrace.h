// dtrace.h — two-level BVH traversal on the device.
//
// Per-lane DFS with the reference's visiting order, so that the winner of
// equal-t hits (t <= t_best accepts a later triangle, geometry/primitives.h:48)
// is the same triangle the CPU picks:
//   BLAS (geometry/blas.cpp:211-307): the root is not box-tested; interior
//     nodes pair-test both children with SSE MINPS/MAXPS semantics; closest
//     hit visits the child with the smaller entry distance first, ties visit
//     the right child first; any-hit visits the left child first.
//   TLAS (geometry/tlas.cpp:172-290): test-on-pop sign-selected slab test;
//     closest hit orders children by their box centres on the parent's
//     longest axis; instances transform the ray without normalising it so t
//     is shared by both spaces (geometry/transform.h:70-77).
//
// Stack: the top `lds_depth` entries of each lane live in LDS, laid out
// entry-major ([entry][lane]) so lanes at equal depth hit distinct banks;
// deeper entries spill, in order, to a per-lane HBM region. The BLAS walk of
// an instance runs on top of the TLAS entries, so both share kMaxStack.
#pragma once
#include "../common/dev_types.h"
#include "dmath.h"

namespace xnh {

struct TravStack {
  uint32_t* lds;        // this lane's column: lds[e * lds_stride]
  uint32_t* spill;      // this lane's spill column: spill[(e - lds_depth) * spill_stride]
  int lds_stride, spill_stride, lds_depth, sp;
  bool overflow;
  __device__ __forceinline__ void push(uint32_t v) {
    if (sp >= kMaxStack) {
      overflow = true;
      return;
    }
    if (sp < lds_depth) lds[sp * lds_stride] = v;
    else spill[(sp - lds_depth) * spill_stride] = v;
    ++sp;
  }
  __device__ __forceinline__ uint32_t pop() {
    --sp;
    return (sp < lds_depth) ? lds[sp * lds_stride] : spill[(sp - lds_depth) * spill_stride];
  }
};

struct TravCount {
  uint32_t node_tests, interior, tri_tests, inst_visits;
};

struct HitRef {
  float t, u, v;
  uint32_t slot;   // global leaf slot
  int32_t inst;    // instance index, -1 = none
};

// Moller-Trumbore on a pre-gathered triangle (geometry/primitives.h:25-49).
__device__ __forceinline__ bool tri_test(const float4* __restrict__ tris, uint32_t slot, f3 o, f3 d, float tmin,
                                         float tmax, float& t, float& u, float& v) {
  const float4 a = tris[3 * slot], b = tris[3 * slot + 1], c = tris[3 * slot + 2];
  const f3 v0{a.x, a.y, a.z}, e1{a.w, b.x, b.y}, e2{b.z, b.w, c.x};
  const f3 h = cross(d, e2);
  const float det = dot(e1, h);
  if (fabsf(det) < 1e-7f) return false;
  const float ia = 1.f / det;
  const f3 s = o - v0;
  u = ia * dot(s, h);
  if (u < 0.f || u > 1.f) return false;
  const f3 q = cross(s, e1);
  v = ia * dot(d, q);
  if (v < 0.f || u + v > 1.f) return false;
  t = ia * dot(e2, q);
  return t >= tmin && t <= tmax;
}

// Two-child slab test; returns the hit mask, near distances in tn0/tn1.
__device__ __forceinline__ int pair_test(float4 la, float4 lb, float4 ra, float4 rb, f3 inv, f3 o, float tmin,
                                         float tmax, float& tn0, float& tn1) {
  const float lx0 = (la.x - o.x) * inv.x, lx1 = (la.w - o.x) * inv.x;
  const float ly0 = (la.y - o.y) * inv.y, ly1 = (lb.x - o.y) * inv.y;
  const float lz0 = (la.z - o.z) * inv.z, lz1 = (lb.y - o.z) * inv.z;
  const float rx0 = (ra.x - o.x) * inv.x, rx1 = (ra.w - o.x) * inv.x;
  const float ry0 = (ra.y - o.y) * inv.y, ry1 = (rb.x - o.y) * inv.y;
  const float rz0 = (ra.z - o.z) * inv.z, rz1 = (rb.y - o.z) * inv.z;
  const float ln = xmax(xmax(xmax(xmin(lx0, lx1), xmin(ly0, ly1)), xmin(lz0, lz1)), tmin);
  const float lf = xmin(xmin(xmin(xmax(lx0, lx1), xmax(ly0, ly1)), xmax(lz0, lz1)), tmax);
  const float rn = xmax(xmax(xmax(xmin(rx0, rx1), xmin(ry0, ry1)), xmin(rz0, rz1)), tmin);
  const float rf = xmin(xmin(xmin(xmax(rx0, rx1), xmax(ry0, ry1)), xmax(rz0, rz1)), tmax);
  tn0 = ln;
  tn1 = rn;
  return (ln <= lf ? 1 : 0) | (rn <= rf ? 2 : 0);
}

// BLAS walk for one instance. Closest: updates tbest/hit on closer (or
// equal) hits. Any: returns true at the first accepted triangle.
template <bool kAny>
__device__ __forceinline__ bool walk_blas(const DevScene& S, uint32_t root, f3 o, f3 d, float tmin, float& tbest,
                                          HitRef& hit, TravStack& st, TravCount& c) {
  if (root == 0xFFFFFFFFu) return false;   // mesh without triangles
  const f3 inv{1.f / d.x, 1.f / d.y, 1.f / d.z};
  const int base = st.sp;
  st.push(root);
  bool found = false;
  while (st.sp > base) {
    const uint32_t ni = st.pop();
    const float4 nb = S.bnodes[2 * ni + 1];
    const uint32_t child = __float_as_uint(nb.z), count = __float_as_uint(nb.w);
    if (count > 0) {
      for (uint32_t i = 0; i < count; ++i) {
        float t, u, v;
        ++c.tri_tests;
        if (tri_test(S.tris, child + i, o, d, tmin, tbest, t, u, v)) {
          if (kAny) {
            st.sp = base;
            return true;
          }
          tbest = t;
          hit.t = t;
          hit.u = u;
          hit.v = v;
          hit.slot = child + i;
          found = true;
        }
      }
      continue;
    }
    const float4 la = S.bnodes[2 * child], lb = S.bnodes[2 * child + 1];
    const float4 ra = S.bnodes[2 * child + 2], rb = S.bnodes[2 * child + 3];
    float tn0, tn1;
    ++c.interior;
    c.node_tests += 2;
    const int m = pair_test(la, lb, ra, rb, inv, o, tmin, tbest, tn0, tn1);
    if (kAny) {
      if (m & 2) st.push(child + 1);
      if (m & 1) st.push(child);
    } else {
      if (m == 3) {
        if (tn0 < tn1) {
          st.push(child + 1);
          st.push(child);
        } else {
          st.push(child);
          st.push(child + 1);
        }
      } else if (m) {
        st.push((m & 1) ? child : child + 1);
      }
    }
    if (st.overflow) {
      st.sp = base;
      return found;
    }
  }
  return found;
}

// Instance: world_to_local columns and metadata (see dev_types.h).
struct InstRec {
  f3 c0, c1, c2, c3;
  uint32_t root;
  int32_t id;
  uint32_t tri_base;   // with kFlatBit
  int32_t lmap;
};
__device__ __forceinline__ InstRec load_inst(const float4* inst, int i) {
  const float4 a = inst[4 * i], b = inst[4 * i + 1], c = inst[4 * i + 2], m = inst[4 * i + 3];
  InstRec r;
  r.c0 = f3{a.x, a.y, a.z};
  r.c1 = f3{a.w, b.x, b.y};
  r.c2 = f3{b.z, b.w, c.x};
  r.c3 = f3{c.y, c.z, c.w};
  r.root = __float_as_uint(m.x);
  r.id = __float_as_int(m.y);
  r.tri_base = __float_as_uint(m.z);
  r.lmap = __float_as_int(m.w);
  return r;
}

// Scene walk: TLAS with nested BLAS walks. Closest hit starts from
// t_best = +inf and ignores tmax, as Scene::intersect does with a default
// HitRecord; any hit uses [tmin, tmax].
template <bool kAny>
__device__ __forceinline__ bool walk_scene(const DevScene& S, f3 o, f3 d, float tmin, float tmax, HitRef& hit,
                                           TravStack& st, TravCount& c) {
  hit.inst = -1;
  if (S.num_tnodes == 0) return false;
  const f3 inv{1.f / d.x, 1.f / d.y, 1.f / d.z};
  const int sx = d.x < 0.f, sy = d.y < 0.f, sz = d.z < 0.f;
  float tbest = kAny ? tmax : __builtin_huge_valf();
  st.sp = 0;
  st.push(0u);
  bool found = false;
  while (st.sp > 0) {
    const uint32_t ni = st.pop();
    const float4 na = S.tnodes[3 * ni], nb = S.tnodes[3 * ni + 1];
    ++c.node_tests;
    // sign-selected slab (geometry/aabb.h:59-71): near/far by direction sign
    const float bx0 = sx ? na.w : na.x, bx1 = sx ? na.x : na.w;
    const float by0 = sy ? nb.x : na.y, by1 = sy ? na.y : nb.x;
    const float bz0 = sz ? nb.y : na.z, bz1 = sz ? na.z : nb.y;
    const float txn = (bx0 - o.x) * inv.x, txf = (bx1 - o.x) * inv.x;
    const float tyn = (by0 - o.y) * inv.y, tyf = (by1 - o.y) * inv.y;
    const float tzn = (bz0 - o.z) * inv.z, tzf = (bz1 - o.z) * inv.z;
    const float lo = smax(smax(smax(tmin, txn), tyn), tzn);
    const float hi = smin(smin(smin(tbest, txf), tyf), tzf);
    if (!(lo <= hi)) continue;
    const uint32_t child = __float_as_uint(nb.z), count = __float_as_uint(nb.w);
    if (count > 0) {
      for (uint32_t k = 0; k < count; ++k) {
        const int ii = static_cast<int>(S.torder[child + k]);
        const InstRec in = load_inst(S.inst, ii);
        ++c.inst_visits;
        const f3 lo_ = in.c0 * o.x + in.c1 * o.y + in.c2 * o.z + in.c3;
        const f3 ld = in.c0 * d.x + in.c1 * d.y + in.c2 * d.z;
        if (kAny) {
          if (walk_blas<true>(S, in.root, lo_, ld, tmin, tbest, hit, st, c)) return true;
        } else {
          if (walk_blas<false>(S, in.root, lo_, ld, tmin, tbest, hit, st, c)) {
            hit.inst = ii;
            found = true;
          }
        }
        if (st.overflow) return found;
      }
    } else {
      ++c.interior;
      const float4 nc = S.tnodes[3 * ni + 2];
      const int ax = static_cast<int>(__float_as_uint(nc.z));
      const float oa = comp(o, ax), ia = comp(inv, ax);
      const float tl = (nc.x - oa) * ia;
      const float tr = (nc.y - oa) * ia;
      if (kAny) {
        st.push(child);
        st.push(child + 1);
      } else if (tl < tr) {
        st.push(child + 1);
        st.push(child);
      } else {
        st.push(child);
        st.push(child + 1);
      }
      if (st.overflow) return found;
    }
  }
  return found;
}

}  // namespace xnh
