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
// BLAS stack entries carry what the walk needs about a node without loading
// it: the node's child index and triangle count. The upload precomputes each
// node's entry into its record, so a pair test reads both children's entries
// with their boxes. Bits 28..31 hold the count (0 = interior: low bits = the
// left child of the pair; 1..14 = leaf: low bits = its first triangle slot;
// 15 = a leaf of 15 or more: low bits = its first slot, whose bigcnt word
// holds the count). Node and slot indices are < 2^28.
//
// Stack: the top `lds_depth` entries of each lane live in LDS, laid out
// entry-major ([entry][lane]) so lanes at equal depth hit distinct banks;
// deeper entries spill, in order, to a per-lane HBM region. The BLAS walk of
// an instance runs on top of the TLAS entries, so both share kMaxStack.
#pragma once
#include "../common/dev_types.h"
#include "dmath.h"

namespace xnh {

// The kernels' dynamic LDS: lds_depth x blockDim.x stack entries.
extern __shared__ uint32_t trav_lds[];

struct TravStack {
  int lds_col;          // this lane's column: trav_lds[e * lds_stride + lds_col]
  int lds_off;          // sp * lds_stride + lds_col (kept as a running index)
  uint32_t* spill;      // this lane's spill column: spill[(e - lds_depth) * spill_stride]
  int lds_stride, spill_stride, lds_depth, sp;   // lds_depth >= 1
  bool overflow;
  __device__ __forceinline__ void reset() {
    sp = 0;
    lds_off = lds_col;
    overflow = false;
  }
  __device__ __forceinline__ void push(uint32_t v) {
    if (sp >= kMaxStack) {
      overflow = true;
      return;
    }
    if (sp < lds_depth) trav_lds[lds_off] = v;
    else gst1(spill, static_cast<size_t>(sp - lds_depth) * spill_stride, v);
    ++sp;
    lds_off += lds_stride;
  }
  // The LDS read is issued unconditionally (at this lane's entry 0 when the
  // entry lies in the spill) so that the common case compiles to a plain
  // LDS access rather than a generic-address load.
  __device__ __forceinline__ uint32_t pop() {
    --sp;
    lds_off -= lds_stride;
    const bool in_lds = sp < lds_depth;
    uint32_t v = trav_lds[in_lds ? lds_off : lds_col];
    if (!in_lds) v = gld1(spill, static_cast<size_t>(sp - lds_depth) * spill_stride);
    return v;
  }
};

// Traversal work counts. The rendering kernels have a counting variant
// (the same walk with these increments) used when work counts are asked
// for; the timed variant carries none of them.
struct TravCount {
  uint32_t node_tests, interior, tri_tests, inst_visits;
  // SIMD use: int_iters, leaf_iters and round_iters count wave iterations
  // (on one lane of the wave); leaf_steps and round_lanes count lanes
  uint32_t int_iters, leaf_iters, leaf_steps, round_iters, round_lanes;
};

// True on the lowest active lane of the wave.
__device__ __forceinline__ bool first_active_lane() {
  return static_cast<int>(threadIdx.x & 63u) == __ffsll(static_cast<long long>(__ballot(1))) - 1;
}

struct HitRef {
  float t, u, v;
  uint32_t slot;   // global leaf slot
  int32_t inst;    // instance index, -1 = none
};

// Moller-Trumbore on a pre-gathered triangle (geometry/primitives.h:25-49).
__device__ __forceinline__ bool tri_test(const float4* __restrict__ tris, uint32_t slot, f3 o, f3 d, float tmin,
                                         float tmax, float& t, float& u, float& v) {
  const float4 a = gld4(tris, 3 * slot), b = gld4(tris, 3 * slot + 1), c = gld4(tris, 3 * slot + 2);
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

// pair_test with the hardware min/max (v_min3/v_max3 chains instead of
// compare-and-select pairs). Equal to pair_test whenever no slab distance is
// NaN, and a NaN needs 0 * inf, so it is used only for rays whose origin,
// direction and inverse direction are finite in every component (box
// coordinates are finite). Without NaNs the two differ at
// most in the sign of a zero distance, which no comparison distinguishes.
__device__ __forceinline__ int pair_test_fast(float4 la, float4 lb, float4 ra, float4 rb, f3 inv, f3 o, float tmin,
                                              float tmax, float& tn0, float& tn1) {
  const float lx0 = (la.x - o.x) * inv.x, lx1 = (la.w - o.x) * inv.x;
  const float ly0 = (la.y - o.y) * inv.y, ly1 = (lb.x - o.y) * inv.y;
  const float lz0 = (la.z - o.z) * inv.z, lz1 = (lb.y - o.z) * inv.z;
  const float rx0 = (ra.x - o.x) * inv.x, rx1 = (ra.w - o.x) * inv.x;
  const float ry0 = (ra.y - o.y) * inv.y, ry1 = (rb.x - o.y) * inv.y;
  const float rz0 = (ra.z - o.z) * inv.z, rz1 = (rb.y - o.z) * inv.z;
  const float ln = fmaxf(fmaxf(fmaxf(fminf(lx0, lx1), fminf(ly0, ly1)), fminf(lz0, lz1)), tmin);
  const float lf = fminf(fminf(fminf(fmaxf(lx0, lx1), fmaxf(ly0, ly1)), fmaxf(lz0, lz1)), tmax);
  const float rn = fmaxf(fmaxf(fmaxf(fminf(rx0, rx1), fminf(ry0, ry1)), fminf(rz0, rz1)), tmin);
  const float rf = fminf(fminf(fminf(fmaxf(rx0, rx1), fmaxf(ry0, ry1)), fmaxf(rz0, rz1)), tmax);
  tn0 = ln;
  tn1 = rn;
  return (ln <= lf ? 1 : 0) | (rn <= rf ? 2 : 0);
}

__device__ __forceinline__ bool all_finite(f3 v) {
  return fabsf(v.x) < __builtin_huge_valf() && fabsf(v.y) < __builtin_huge_valf() &&
         fabsf(v.z) < __builtin_huge_valf();
}

// Written with selects rather than branches: it runs for both children of
// every interior step.
// Instance: world_to_local columns and metadata (see dev_types.h).
struct InstRec {
  f3 c0, c1, c2, c3;
  uint32_t root;
  int32_t id;
  uint32_t tri_base;   // with kFlatBit
  int32_t lmap;
};
__device__ __forceinline__ InstRec load_inst(const float4* inst, int i) {
  const float4 a = gld4(inst, 4 * i), b = gld4(inst, 4 * i + 1), c = gld4(inst, 4 * i + 2), m = gld4(inst, 4 * i + 3);
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

// ---- analytic primitives ------------------------------------------------------
// geometry/primitives.h:54-318 and geometry/prim_blas.cpp:165-229, with the
// reference's operation order (true divisions, std::min/std::max and the
// first-largest rule of std::max({...})).
struct PrimRec {
  int type, mat;
  float q[10];
};
__device__ __forceinline__ PrimRec load_prim(const float4* P, uint32_t i) {
  const float4 a = gld4(P, 3 * static_cast<size_t>(i)), b = gld4(P, 3 * static_cast<size_t>(i) + 1),
               c = gld4(P, 3 * static_cast<size_t>(i) + 2);
  PrimRec r;
  r.type = __float_as_int(a.x);
  r.mat = __float_as_int(a.y);
  r.q[0] = a.z; r.q[1] = a.w; r.q[2] = b.x; r.q[3] = b.y; r.q[4] = b.z;
  r.q[5] = b.w; r.q[6] = c.x; r.q[7] = c.y; r.q[8] = c.z; r.q[9] = c.w;
  return r;
}

// AABB::intersect (geometry/aabb.h:34-47) on a PrimBLAS node.
__device__ __forceinline__ bool prim_node_hit(float4 na, float4 nb, f3 o, f3 d, float rtmin, float rtmax,
                                              float& tnear) {
  const float tx0 = (na.x - o.x) / d.x, tx1 = (na.w - o.x) / d.x;
  const float ty0 = (na.y - o.y) / d.y, ty1 = (nb.x - o.y) / d.y;
  const float tz0 = (na.z - o.z) / d.z, tz1 = (nb.y - o.z) / d.z;
  float lo = smin(tx0, tx1), hi = smax(tx0, tx1);
  const float ly = smin(ty0, ty1), lz = smin(tz0, tz1), hy = smax(ty0, ty1), hz = smax(tz0, tz1);
  if (lo < ly) lo = ly;
  if (lo < lz) lo = lz;
  if (lo < rtmin) lo = rtmin;
  if (hy < hi) hi = hy;
  if (hz < hi) hi = hz;
  if (rtmax < hi) hi = rtmax;
  tnear = lo;
  return lo <= hi;
}

// Box slab walk (primitives.h:159-182): the entry distance and the entering
// face (axis, and whether it is the low face).
__device__ __forceinline__ bool prim_box_slabs(const PrimRec& p, f3 o, f3 d, float tmin, float rect, float& t,
                                               int& axis, bool& low) {
  const f3 c{p.q[0], p.q[1], p.q[2]}, h{p.q[3], p.q[4], p.q[5]};
  const f3 mn = c - h, mx = c + h;
  float t_lo = tmin, t_hi = rect;
  axis = 0;
  low = false;
  for (int i = 0; i < 3; ++i) {
    const float inv = 1.f / comp(d, i);
    float t0 = (comp(mn, i) - comp(o, i)) * inv;
    float t1 = (comp(mx, i) - comp(o, i)) * inv;
    const bool flip = inv < 0.f;
    if (flip) {
      const float tmp = t0;
      t0 = t1;
      t1 = tmp;
    }
    if (t0 > t_lo) {
      t_lo = t0;
      axis = i;
      low = !flip;
    }
    t_hi = smin(t_hi, t1);
    if (t_lo > t_hi) return false;
  }
  t = t_lo;
  return true;
}

// The primitive's intersect: a hit at t in [tmin, rect).
__device__ __forceinline__ bool prim_hit(const PrimRec& p, f3 o, f3 d, float tmin, float rect, float& t) {
  switch (p.type) {
    case 0: {   // sphere
      const f3 c{p.q[0], p.q[1], p.q[2]};
      const float r = p.q[3];
      const f3 oc = o - c;
      const float a = dot(d, d), hb = dot(oc, d), cc = dot(oc, oc) - r * r;
      const float disc = hb * hb - a * cc;
      if (disc < 0.f) return false;
      const float sq = sqrtf(disc);
      t = (-hb - sq) / a;
      if (t < tmin || t >= rect) {
        t = (-hb + sq) / a;
        if (t < tmin || t >= rect) return false;
      }
      return true;
    }
    case 1: {   // plane
      const f3 n{p.q[0], p.q[1], p.q[2]};
      const float denom = dot(n, d);
      if (fabsf(denom) < 1e-7f) return false;
      t = (p.q[3] - dot(n, o)) / denom;
      return !(t < tmin || t >= rect);
    }
    case 2: {   // box
      int axis;
      bool low;
      return prim_box_slabs(p, o, d, tmin, rect, t, axis, low);
    }
    case 3: {   // disk
      const f3 c{p.q[0], p.q[1], p.q[2]}, n{p.q[3], p.q[4], p.q[5]};
      const float denom = dot(n, d);
      if (fabsf(denom) < 1e-7f) return false;
      t = dot(c - o, n) / denom;
      if (t < tmin || t >= rect) return false;
      const f3 op = (o + t * d) - c;
      return !(dot(op, op) > p.q[6] * p.q[6]);
    }
    default: {   // quad
      const f3 org{p.q[0], p.q[1], p.q[2]}, u{p.q[3], p.q[4], p.q[5]}, v{p.q[6], p.q[7], p.q[8]};
      const f3 nr = cross(u, v);
      const float area_sq = dot(nr, nr);
      if (area_sq < 1e-14f) return false;
      const float denom = dot(nr, d);
      if (fabsf(denom) < 1e-9f) return false;
      t = dot(nr, org - o) / denom;
      if (t < tmin || t >= rect) return false;
      const f3 op = (o + t * d) - org;
      const float s = dot(cross(op, v), nr) / area_sq;
      const float tv = dot(cross(u, op), nr) / area_sq;
      return !(s < 0.f || s > 1.f || tv < 0.f || tv > 1.f);
    }
  }
}

// Normal (facing the ray) and (u, v) the primitive's intersect writes for a
// hit at t.
__device__ __forceinline__ f3 prim_normal(const PrimRec& p, f3 o, f3 d, float tmin, float t, float& u, float& v) {
  constexpr float kPi = 3.14159265358979323846f;
  constexpr float kInvPi = 1.f / kPi;
  const f3 pos = o + t * d;
  switch (p.type) {
    case 0: {
      const f3 c{p.q[0], p.q[1], p.q[2]};
      const f3 out = (pos - c) / p.q[3];
      const float phi = atan2f(-out.z, out.x) + kPi;
      const float theta = acosf(fminf(fmaxf(-out.y, -1.f), 1.f));
      u = phi * (0.5f * kInvPi);
      v = theta * kInvPi;
      return dot(out, d) < 0.f ? out : -out;
    }
    case 1: {
      const f3 n{p.q[0], p.q[1], p.q[2]};
      u = pos.x - floorf(pos.x);
      v = pos.z - floorf(pos.z);
      return dot(n, d) < 0.f ? n : -n;
    }
    case 2: {
      float tt;
      int axis;
      bool low;
      prim_box_slabs(p, o, d, tmin, __builtin_huge_valf(), tt, axis, low);
      f3 n{0.f, 0.f, 0.f};
      const float s = low ? -1.f : 1.f;
      if (axis == 0) n.x = s;
      else if (axis == 1) n.y = s;
      else n.z = s;
      if (dot(n, d) > 0.f) n = -n;
      const f3 c{p.q[0], p.q[1], p.q[2]}, h{p.q[3], p.q[4], p.q[5]};
      const f3 mn = c - h;
      const int ua = (axis + 1) % 3, va = (axis + 2) % 3;
      u = (comp(pos, ua) - comp(mn, ua)) / (2.f * comp(h, ua));
      v = (comp(pos, va) - comp(mn, va)) / (2.f * comp(h, va));
      return n;
    }
    case 3: {
      const f3 c{p.q[0], p.q[1], p.q[2]}, n{p.q[3], p.q[4], p.q[5]};
      const f3 op = pos - c;
      const f3 ref = fabsf(n.x) < 0.9f ? f3{1.f, 0.f, 0.f} : f3{0.f, 1.f, 0.f};
      const f3 tang = normalize(cross(n, ref));
      const f3 bita = cross(n, tang);
      float phi = atan2f(dot(op, bita), dot(op, tang));
      if (phi < 0.f) phi += 2.f * kPi;
      u = phi * (0.5f * kInvPi);
      v = sqrtf(dot(op, op)) / p.q[6];
      return dot(n, d) < 0.f ? n : -n;
    }
    default: {
      const f3 org{p.q[0], p.q[1], p.q[2]}, uu{p.q[3], p.q[4], p.q[5]}, vv{p.q[6], p.q[7], p.q[8]};
      const f3 nr = cross(uu, vv);
      const float area_sq = dot(nr, nr);
      const f3 op = pos - org;
      u = dot(cross(op, vv), nr) / area_sq;
      v = dot(cross(uu, op), nr) / area_sq;
      const f3 n = normalize(nr);
      return dot(nr, d) > 0.f ? -n : n;
    }
  }
}

// Scene walk as a resumable state machine, advanced one round at a time so
// a persistent kernel can retire a finished ray and start the next in the
// same lane while the rest of the wave keeps walking.
//
// A round runs the lane's TLAS steps until it enters a BLAS walk, then two
// wave-wide phases over the lanes inside BLAS walks (the while-while loop of
// Aila and Laine): an interior phase that walks node pairs, continuing with
// the nearer hit child in registers and pushing only the farther, until the
// lane reaches a leaf (held in a slot), and a leaf phase that tests the held
// leaves, continuing through consecutive leaves. Each phase runs one kind
// of step for all its lanes instead of both kinds every iteration.
//
// No step is taken early: a lane holding a leaf walks no further interior
// nodes until the leaf is tested, since a node box tested against a t_best
// not yet lowered by the leaf's hits may let through a subtree the reference
// culls, and box and triangle rounding differ enough for such a subtree to
// hold a hit. Every box and triangle test therefore sees the reference's
// t_best, and results and work counts equal its recursive walk, ties
// included. Closest-hit rays follow the reference's visiting order: BLAS
// children nearer entry first, ties right first; TLAS test-on-pop with
// centre-ordered children (tlas.cpp:231-241); a TLAS leaf's instances
// (pushed with kInstTag, in reverse) walked first to last. Any-hit rays
// take BLAS children farther first (the reference: left first,
// blas.cpp:303-304); the occlusion answer does not depend on the order, only
// the work does. Closest hit starts from t_best = +inf and
// ignores tmax, as Scene::intersect does with a default HitRecord; any hit
// uses [tmin, tmax]. A stack overflow ends the ray with what it has found.
constexpr uint32_t kInstTag = 0x80000000u;
constexpr uint32_t kEntEnd = kNoRoot;   // no entry left in the BLAS walk

__device__ __forceinline__ bool ent_interior(uint32_t e) { return e != kEntEnd && (e >> kEntShift) == 0u; }
__device__ __forceinline__ bool ent_leaf(uint32_t e) { return e != kEntEnd && (e >> kEntShift) != 0u; }

// kPrims: the scene may hold analytic primitives (a variant, because their
// walk raises the register count of the whole kernel).
template <bool kAny, bool kCount = true, bool kPrims = true>
struct Walker {
  f3 o, d, inv;       // world ray
  f3 lo, ld, linv;    // instance-space ray of the BLAS walk in progress
  float tmin, tbest;
  float rtmax;        // the ray's tmax (bounds the primitives' box tests, aabb.h:43)
  int blas_base;      // stack depth under the current BLAS walk; -1 in the TLAS
  int inst;           // instance of the BLAS walk
  uint32_t e;         // next BLAS entry to process, or kEntEnd; a leaf only while `leaf` is held
  uint32_t leaf;      // held leaf entry (tested before e), 0 = none
  bool inst_found, found;
  bool fast;          // lo, ld, linv finite: pair_test_fast is exact
  bool stuck;         // the step bound was reached (the walk was abandoned)
  HitRef hit;         // u, v, slot, inst of the closest hit; its t is tbest

  __device__ __forceinline__ void begin(const DevScene& S, f3 o_, f3 d_, float tmin_, float tmax_, TravStack& st) {
    o = o_;
    d = d_;
    inv = f3{1.f / d.x, 1.f / d.y, 1.f / d.z};
    tmin = tmin_;
    rtmax = tmax_;
    tbest = kAny ? tmax_ : __builtin_huge_valf();
    blas_base = -1;
    inst = -1;
    e = kEntEnd;
    leaf = 0u;
    inst_found = false;
    found = false;
    fast = false;
    stuck = false;
    hit = HitRef{0.f, 0.f, 0.f, 0u, -1};
    st.reset();
    if (S.num_tnodes > 0) st.push(0u);
  }

  __device__ __forceinline__ void end_instance() {
    if (!kAny && inst_found) {
      hit.inst = inst;
      found = true;
    }
    blas_base = -1;
  }

  __device__ __forceinline__ uint32_t pop_blas(TravStack& st) { return st.sp > blas_base ? st.pop() : kEntEnd; }

  // The scene's primitives, walked to the end at once in the instance-space
  // ray lo/ld (PrimBLAS::intersect / ::intersects, prim_blas.cpp:165-229):
  // nodes tested when popped, children tested (closest hit) and pushed
  // nearer last; true when an any-hit walk found an occluder.
  __device__ __forceinline__ bool prim_walk(const DevScene& S, TravStack& st, TravCount& c) {
    if (S.num_pnodes <= 0) return false;
    const int base = st.sp;
    st.push(0u);
    while (st.sp > base && !st.overflow) {
      const uint32_t idx = st.pop();
      const float4 na = gld4(S.pnodes, 2 * static_cast<size_t>(idx)), nb = gld4(S.pnodes, 2 * static_cast<size_t>(idx) + 1);
      float tn;
      if (kCount) ++c.node_tests;
      if (!prim_node_hit(na, nb, lo, ld, tmin, rtmax, tn)) continue;
      if (!kAny && tn > tbest) continue;
      const uint32_t child = __float_as_uint(nb.z), count = __float_as_uint(nb.w);
      if (count > 0) {
        for (uint32_t k = 0; k < count; ++k) {
          const uint32_t pi = gld1(S.porder, child + k);
          const PrimRec pr = load_prim(S.prims, pi);
          float t;
          if (kCount) ++c.tri_tests;
          if (prim_hit(pr, lo, ld, tmin, kAny ? rtmax : tbest, t)) {
            if (kAny) return true;   // the ray is finished; its stack is reset by the next begin()
            tbest = t;
            hit.u = 0.f;
            hit.v = 0.f;
            hit.slot = pi | kPrimHit;
            inst_found = true;
          }
        }
        continue;
      }
      if (kAny) {
        st.push(child);
        st.push(child + 1);
        continue;
      }
      if (kCount) {
        ++c.interior;
        c.node_tests += 2;
      }
      float tl, tr;
      const bool hl = prim_node_hit(gld4(S.pnodes, 2 * static_cast<size_t>(child)),
                                    gld4(S.pnodes, 2 * static_cast<size_t>(child) + 1), lo, ld, tmin, rtmax, tl) &&
                      tl <= tbest;
      const bool hr = prim_node_hit(gld4(S.pnodes, 2 * static_cast<size_t>(child) + 2),
                                    gld4(S.pnodes, 2 * static_cast<size_t>(child) + 3), lo, ld, tmin, rtmax, tr) &&
                      tr <= tbest;
      if (hl && hr) {
        if (tl < tr) {
          st.push(child + 1);
          st.push(child);
        } else {
          st.push(child);
          st.push(child + 1);
        }
      } else if (hl) {
        st.push(child);
      } else if (hr) {
        st.push(child + 1);
      }
    }
    return false;
  }

  // One TLAS step: pops one TLAS entry, a node (tested on pop, its children
  // or instances pushed) or an instance to enter (its BLAS walk starts, or
  // the primitives are walked to the end). True once the ray is finished.
  __device__ __forceinline__ bool tlas_step(const DevScene& S, TravStack& st, TravCount& c) {
    if (st.sp == 0) return true;
    const uint32_t te = st.pop();
    if (te & kInstTag) {
      const int ii = static_cast<int>(te & ~kInstTag);
      const InstRec in = load_inst(S.inst, ii);
      if (kCount) ++c.inst_visits;
      lo = in.c0 * o.x + in.c1 * o.y + in.c2 * o.z + in.c3;
      ld = in.c0 * d.x + in.c1 * d.y + in.c2 * d.z;
      linv = f3{1.f / ld.x, 1.f / ld.y, 1.f / ld.z};
      fast = all_finite(linv) && all_finite(ld) && all_finite(lo);
      inst = ii;
      inst_found = false;
      if (in.root == kPrimRoot) {   // the primitives: walked here, to the end
        if (kPrims) {
          if (prim_walk(S, st, c)) {
            found = true;
            return true;
          }
          if (!kAny && inst_found) {
            hit.inst = ii;
            found = true;
          }
          if (st.overflow) return true;
        }
        return false;
      }
      blas_base = st.sp;
      // e never holds a leaf unless one is already held: a root that is a
      // leaf is held at once (kNoRoot, an empty mesh, is kEntEnd)
      e = in.root;
      leaf = 0u;
      if (ent_leaf(e)) {
        leaf = e;
        e = kEntEnd;
      }
      return false;
    }
    const uint32_t ni = te;
    const float4 na = gld4(S.tnodes, 3 * ni), nb = gld4(S.tnodes, 3 * ni + 1);
    if (kCount) ++c.node_tests;
    // sign-selected slab (geometry/aabb.h:59-71): near/far by direction sign
    const bool sx = d.x < 0.f, sy = d.y < 0.f, sz = d.z < 0.f;
    const float bx0 = sx ? na.w : na.x, bx1 = sx ? na.x : na.w;
    const float by0 = sy ? nb.x : na.y, by1 = sy ? na.y : nb.x;
    const float bz0 = sz ? nb.y : na.z, bz1 = sz ? na.z : nb.y;
    const float txn = (bx0 - o.x) * inv.x, txf = (bx1 - o.x) * inv.x;
    const float tyn = (by0 - o.y) * inv.y, tyf = (by1 - o.y) * inv.y;
    const float tzn = (bz0 - o.z) * inv.z, tzf = (bz1 - o.z) * inv.z;
    const float lo_t = smax(smax(smax(tmin, txn), tyn), tzn);
    const float hi_t = smin(smin(smin(tbest, txf), tyf), tzf);
    if (!(lo_t <= hi_t)) return false;
    const uint32_t child = __float_as_uint(nb.z), count = __float_as_uint(nb.w);
    if (count > 0) {
      for (uint32_t k = count; k-- > 0;) st.push(kInstTag | gld1(S.torder, child + k));
    } else {
      if (kCount) ++c.interior;
      const float4 nc = gld4(S.tnodes, 3 * ni + 2);
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
    }
    return st.overflow;
  }

  // The kind of step the lane needs next: 0 a TLAS step, 1 the held leaf,
  // 2 an interior step. A finished BLAS walk (no entry left, no leaf held)
  // first returns the lane to the TLAS.
  __device__ __forceinline__ int next_kind() {
    if (blas_base >= 0 && leaf == 0u && !ent_interior(e)) end_instance();
    return blas_base < 0 ? 0 : (leaf != 0u ? 1 : 2);
  }

  // One interior step on e (an interior entry, no leaf held): pair-test its
  // children, continue with the nearer hit one, push the farther; a leaf
  // reached is held and the next entry popped (walked after the leaf).
  template <bool kFast>
  __device__ __forceinline__ void interior_step(const DevScene& S, TravStack& st, TravCount& c) {
    const uint32_t child = e;
    const float4 la = gld4(S.bnodes, 2 * child), lb = gld4(S.bnodes, 2 * child + 1);
    const float4 ra = gld4(S.bnodes, 2 * child + 2), rb = gld4(S.bnodes, 2 * child + 3);
    float tn0, tn1;
    if (kCount) {
      ++c.interior;
      c.node_tests += 2;
    }
    const int m = kFast ? pair_test_fast(la, lb, ra, rb, linv, lo, tmin, tbest, tn0, tn1)
                        : pair_test(la, lb, ra, rb, linv, lo, tmin, tbest, tn0, tn1);
    const uint32_t el = __float_as_uint(lb.w), er = __float_as_uint(rb.w);
    uint32_t next;
    if (m == 3) {
      // closest: nearer first, ties right first (blas.cpp:252-258). Any
      // hit: farther first. Occlusion does not depend on the order (every
      // box is tested against the fixed tmax); a shadow ray's nearer child
      // is mostly the box around its own origin, whose triangles it cannot
      // hit, so the farther child meets occluders sooner.
      const bool left_first = kAny ? !(tn0 < tn1) : tn0 < tn1;
      st.push(left_first ? er : el);
      next = left_first ? el : er;
    } else if (m) {
      next = (m & 1) ? el : er;
    } else {
      next = pop_blas(st);
    }
    if (ent_leaf(next)) {
      leaf = next;
      next = pop_blas(st);
    }
    e = next;
  }

  // Tests the held leaf; true when an any-hit walk found an occluder.
  __device__ __forceinline__ bool leaf_step(const DevScene& S, TravStack& st, TravCount& c) {
    uint32_t count = leaf >> kEntShift;
    const uint32_t first = leaf & kEntIndexMask;
    if (count == kEntBigLeaf) count = gld1(S.bigcnt, first);
    for (uint32_t k = 0; k < count; ++k) {
      float t, u, v;
      if (kCount) ++c.tri_tests;
      if (tri_test(S.tris, first + k, lo, ld, tmin, tbest, t, u, v)) {
        if (kAny) {
          found = true;
          return true;
        }
        tbest = t;   // the hit distance (hit.t is not kept during the walk)
        hit.u = u;
        hit.v = v;
        hit.slot = first + k;
        inst_found = true;
      }
    }
    leaf = 0u;
    if (ent_leaf(e)) {
      leaf = e;
      e = pop_blas(st);
    }
    return false;
  }

};

// One scheduled run of a wave's walks. Every lane holding a ray names the
// kind of step it needs next (TLAS entry, held leaf, or interior node pair);
// the kind most lanes need runs, on those lanes only, while the others wait:
// interior and leaf steps repeat while their kind keeps the majority of the
// wave's BLAS lanes, TLAS steps while at least half of the lanes that began
// the run still need one. Each lane's own steps still run one at a time in
// the reference's order, each seeing the t_best its earlier steps left:
// only the interleaving of different lanes' steps is chosen here, so results
// and work counts are the reference's. Scheduling by majority rather than
// letting every lane finish a whole descent before any leaf is tested keeps
// more lanes of each iteration busy (a wave's descents differ in length).
// True on lanes whose ray finished (closest: `found` / `hit`; any: `found` =
// occluded). Every lane of the wave must reach it.
// Interior steps of wave_step's lanes while interior steps keep the
// majority of the wave's BLAS lanes.
template <bool kFast, bool kAny, bool kCount, bool kPrims>
__device__ __forceinline__ void interior_run(Walker<kAny, kCount, kPrims>& w, const DevScene& S, TravStack& st,
                                             TravCount& c, int& k, int& ni, int& nl, bool& done) {
  constexpr int kLeaf = 1, kInterior = 2, kWait = 3;
  for (;;) {
    if (kCount && first_active_lane()) ++c.int_iters;
    if (k == kInterior) {
      w.template interior_step<kFast>(S, st, c);
      if (st.overflow) {
        w.end_instance();
        done = true;
        k = -1;
      } else {
        k = w.leaf != 0u ? kLeaf : (ent_interior(w.e) ? kInterior : kWait);
      }
    }
    ni = __popcll(__ballot(k == kInterior));
    nl = __popcll(__ballot(k == kLeaf));
    if (ni == 0 || ni < nl) break;
  }
}

template <bool kAny, bool kCount, bool kPrims>
__device__ __forceinline__ bool wave_step(Walker<kAny, kCount, kPrims>& w, const DevScene& S, TravStack& st,
                                          TravCount& c, bool active) {
  constexpr int kTlas = 0, kLeaf = 1, kInterior = 2, kWait = 3;
  int k = active ? w.next_kind() : -1;
  const int nt = __popcll(__ballot(k == kTlas));
  int nl = __popcll(__ballot(k == kLeaf)), ni = __popcll(__ballot(k == kInterior));
  bool done = false;
  if (ni >= nl && ni >= nt) {
    // the hardware min/max slab test when every walking lane's ray allows it
    // (a wave-uniform choice). The two runs are separate loops: with one
    // loop and the choice inside it, the compiler computed the slab
    // distances before the branch, and the min/max instructions then
    // re-canonicalised each of them (a value from another block may be a
    // signalling NaN), 12 more instructions per node pair.
    if (__ballot(k == kInterior && !w.fast) == 0ull) interior_run<true>(w, S, st, c, k, ni, nl, done);
    else interior_run<false>(w, S, st, c, k, ni, nl, done);
  } else if (nl >= nt) {
    for (;;) {
      if (kCount && first_active_lane()) ++c.leaf_iters;
      if (k == kLeaf) {
        if (kCount) ++c.leaf_steps;
        if (w.leaf_step(S, st, c)) {
          done = true;
          k = -1;
        } else {
          k = w.leaf != 0u ? kLeaf : (ent_interior(w.e) ? kInterior : kWait);
        }
      }
      nl = __popcll(__ballot(k == kLeaf));
      ni = __popcll(__ballot(k == kInterior));
      if (nl == 0 || nl < ni) break;
    }
  } else {
    for (;;) {
      if (k == kTlas) {
        if (w.tlas_step(S, st, c) || st.overflow) {
          if (st.overflow) w.end_instance();
          done = true;
          k = -1;
        } else if (w.blas_base >= 0) {
          k = kWait;
        }
      }
      const int n0 = __popcll(__ballot(k == kTlas));
      if (n0 == 0 || 2 * n0 < nt) break;
    }
  }
  return done;
}

// Whole walk of one ray per lane (the ray-batch API), with work counts.
// Every lane of the wave must reach it.
template <bool kAny>
__device__ __forceinline__ bool walk_scene(const DevScene& S, f3 o, f3 d, float tmin, float tmax, bool valid,
                                           HitRef& hit, TravStack& st, TravCount& c) {
  Walker<kAny> w;
  w.begin(S, o, d, tmin, tmax, st);
  bool done = !valid;
  for (uint32_t guard = 0; __ballot(!done) != 0ull; ++guard) {
    if (guard >= kWaveStepLimit) {
      w.stuck = true;
      break;
    }
    if (wave_step(w, S, st, c, !done)) done = true;
  }
  if (w.stuck) st.overflow = true;
  hit = w.hit;
  hit.t = w.found ? w.tbest : 0.f;
  return valid && w.found;
}

}  // namespace xnh
