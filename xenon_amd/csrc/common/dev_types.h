// dev_types.h — HBM layout of the scene and of the wavefront queues, shared
// by the host API (which fills it) and the kernels (which read it).
//
// Scene (read-only, built once by upload_scene):
//   bnodes  2 x float4 per BLAS node, all meshes concatenated:
//           {lo.x, lo.y, lo.z, hi.x} {hi.y, hi.z, child, count}
//           interior: child = global index of the left child (right = +1);
//           leaf: child = global triangle slot of its first triangle.
//   tris    3 x float4 per leaf slot, in leaf order:
//           {v0.xyz, e1.x} {e1.yz, e2.xy} {e2.z, tri, 0, 0}
//           (e1 = v1 - v0, e2 = v2 - v0, computed once on the host: the same
//           IEEE subtraction the reference performs per test)
//   tnodes  3 x float4 per TLAS node: the two BLAS-style float4 plus
//           {centre of left child, centre of right child, axis, 0} along the
//           node's longest axis, the traversal order key of tlas.cpp:231-241
//   inst    4 x float4 per instance: world_to_local columns c0..c3 packed
//           {c0.xyz, c1.x} {c1.yz, c2.xy} {c2.z, c3.xyz} and
//           {blas root node, instance_id, tri_base | flat<<31, light map or -1}
//   vidx    3 x int per global triangle (global vertex indices)
//   vnrm    float4 per global vertex (xyz normal)
//   tmat    material id per global triangle
//   lmap    per instance whose id names area lights: first light index per
//           mesh triangle, -1 for none
//   lights  4 x float4 per light: {p0, area} {p1, e.r} {p2, e.g} {n_geo, e.b}
//   mats    4 x float4 per material (see DevMat)
//   suns    2 x float4 per directional light: {dir, 0} {emission, 0}
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

namespace xnh {

constexpr int kMaxStack = 128;          // TLAS 64 + BLAS 64, the reference's two stacks
constexpr uint32_t kFlatBit = 0x80000000u;

struct DevSky {
  float zenith[3], horizon[3], ground[3];
  float sharpness, intensity;
};

struct DevScene {
  const float4* bnodes;
  const float4* tris;
  const float4* tnodes;
  const uint32_t* torder;
  const float4* inst;
  const int32_t* vidx;
  const float4* vnrm;
  const int32_t* tmat;
  const int32_t* lmap;
  const float4* lights;
  const float4* mats;
  const float4* suns;
  int32_t num_tnodes;
  int32_t num_lights;
  int32_t num_suns;
  int32_t pad;
  DevSky sky;
};

struct DevCamera {
  float eye[3], llc[3], horiz[3], vert[3];
};

// Wavefront state. Path queues are double-buffered; entry i of a queue is
// one live path. Per-sample ("slot") buffers are indexed by
// slot = sample_in_batch * num_pixels + pixel and hold radiance sums.
struct PathQueue {
  float4* ray0;     // {o.xyz, d.x}
  float4* ray1;     // {d.y, d.z, prev_pdf, flags}
  float4* thr;      // {T.xyz, 0}
  uint4* st;        // {rng lo, rng hi, slot, 0}
};

struct WaveBuffers {
  PathQueue q[2];
  float4* hit;      // {t, u, v, instance index (bits), } per queue entry; inst -1 = miss
  uint32_t* hit_slot;  // global triangle slot per queue entry
  float4* sh0;      // shadow ray {o.xyz, d.x}
  float4* sh1;      // {d.y, d.z, tmax, slot*2+k}
  float4* sh_c;     // unoccluded contribution
  float4* l_term;   // per slot: emission / sky at path end
  float4* l_nee;    // per slot x 2: area and sun NEE sums
  float4* accum;    // per pixel running sum
  uint32_t* counters;   // see Ctr
  unsigned long long* stats;  // see Stat
  uint32_t* spill;  // traversal stack entries beyond the LDS part
  int32_t capacity;
};

enum Ctr : int {
  kCount0 = 0,        // live paths in queue 0
  kCount1 = 1,        // live paths in queue 1
  kShadowCount = 2,
  kOverflow = 3,      // lanes whose stack exceeded kMaxStack
  kCursorBase = 8,    // dynamic work cursors, one per launch of a batch
  kNumCtr = 8 + 4 * 64
};

enum Stat : int {
  kStPaths = 0,
  kStClosest,
  kStShadow,
  kStNodeTests,
  kStInterior,
  kStTriTests,
  kStInstVisits,
  kStShNodeTests,     // the same four for any-hit (shadow) traversal
  kStShInterior,
  kStShTriTests,
  kStShInstVisits,
  kStNonFinite,
  kNumStat
};

}  // namespace xnh
