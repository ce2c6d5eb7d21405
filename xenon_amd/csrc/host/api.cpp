// api.cpp — renderer half of the C ABI of libxenon_hip.so
// (include/xenon_hip.h).
//
// Device selection, buffer ownership, scene flattening into the HBM layout
// of common/dev_types.h, batch scheduling and the bounce loop. A render call
// is a list of jobs; their samples are cut into batches of chunks (one
// contiguous sample range of one job per chunk), all chunk tables are
// uploaded once, and each batch runs
//
//   raygen -> [closest(d) -> shade(d) -> shadow(d)] d < max depth -> accumulate -> advance
//
// with no host read between kernels: every work count and the batch index
// live in device memory, so the sequence is captured once as a hipGraph and
// replayed per batch. With per-kernel profiling on, the same sequence is
// launched directly with HIP events around the traversal and shading
// kernels.
//
// The reference's CUDA backend aborts on any CUDA error
// (cuda/cuda_utils.cuh:12-20) and reads counters back ~6 times per bounce
// (cuda/cuda_renderer.cu:516-573); neither is carried over.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../common/dev_types.h"
#include "../common/guide_net.h"
#include "../device/launch.h"
#include "xenon_hip.h"

using namespace xnh;

struct xnh_gpu_scene {
  std::vector<void*> allocs;
  DevScene S{};
  size_t bytes = 0;
};

// Path-guiding network state: FP32 masters, FP16 copies, gradients, Adam
// moments, and the device training batch.
struct GuideState {
  bool on = false;
  xnh_guide_config cfg{};
  GuideNet net{};
  std::vector<void*> allocs;
  float *grid_m = nullptr, *grid_g = nullptr, *grid_am = nullptr, *grid_av = nullptr;
  uint16_t* grid_h = nullptr;
  float *w_m = nullptr, *w_g = nullptr, *w_am = nullptr, *w_av = nullptr;
  uint16_t* w_h = nullptr;
  uint16_t* w_t = nullptr;        // FP16 weights as MFMA operand tiles
  float* loss = nullptr;
  size_t ngrid = 0;
  int step = 0;
  // guided rendering: per-entry mixtures, training records and the device
  // training batch built from them
  float4* mix = nullptr;
  uint4* feat = nullptr;          // hash-grid encodings of a queue (6 uint4 per entry)
  float4* rec = nullptr;
  float *t_pos = nullptr, *t_nrm = nullptr, *t_dir = nullptr, *t_lam = nullptr, *t_lam_sum = nullptr;
  float *lam_all = nullptr, *lam_sum_all = nullptr;   // lambda of every record, and their sum
  float* w_part = nullptr;        // per-work-group weight gradients of a training launch
  int w_part_blocks = 0;
  uint32_t* t_n = nullptr;
  uint32_t collects = 0;          // training batches drawn (seeds the selection)
  float* alpha_dev = nullptr;     // MIS probability of the guide for the next batch
  float alpha_host = 0.f;
  bool train = true;
  uint32_t mix_stride = 0;        // queue entries of a guided batch (stride of mix and feat)
  uint32_t slot_cap = 0;          // paths (slots) of a guided batch
  uint32_t rec_cap = 0, rec_every = 0;
};

struct xnh_ctx {
  xnh_config cfg{};
  int device = 0;                 // the device current at xnh_create
  GuideState guide;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  std::vector<hipEvent_t> prof_events;
  bool profiling = false;         // per-kernel HIP-event timing
  bool counting = false;          // traversal work counters (the counting kernel variants)
  bool prims = false;             // a scene of the current jobs holds analytic primitives
  bool use_graph = true;
  xnh_gpu_scene own;              // the scene of xnh_upload_scene
  bool has_scene = false;
  WaveBuffers B{};                // B.capacity 0: no queues (a re-tuning failed twice)
  std::vector<void*> wave_allocs;
  std::vector<void*> pix_allocs;  // accum and d_out: sized by the resolution, kept across re-tunings
  float4* accum = nullptr;        // the contract's running per-pixel sums
  DevJob* d_jobs = nullptr;
  DevChunk* d_chunks = nullptr;
  int32_t* d_nchunks = nullptr;
  size_t jobs_cap = 0, batches_cap = 0;
  uint64_t table_gen = 0;         // bumps when a table is reallocated
  hipGraphExec_t graph = nullptr;
  int graph_depth = -1;
  uint64_t graph_gen = 0;
  int nsamples = 1;               // samples per pixel per batch at the configured resolution
  int block = 256;
  int lds_depth = 16;
  int grid_trace = 0;
  int grid_shade = 0;
  float* d_out = nullptr;
  int spp_total = 0;
  struct {                        // output-stage scratch (xnh_image_noise_stats, xnh_image_rgb8)
    int64_t cap = 0;              // pixels
    float* img = nullptr;         // W*H*3 framebuffer
    uint8_t* rgb8 = nullptr;
    ImgScratch sc{};
    std::vector<void*> allocs;
  } out;
  xnh_stats stats{};
};

namespace {

// Default samples per pixel per batch: about 2^28 paths in flight (about
// 75 GB of queues at 292 B per path, a quarter of the device), at most 256
// samples. Each launch pays a fixed tail (its last, longest walks) and every
// batch the kernel boundaries of all its bounces, so throughput keeps rising
// with the batch: 1080p benchmark, 4 samples per batch (8.3 M paths) 559
// Mpaths/s, 16 samples 834, 64 samples 960 (977 with later changes), 128
// samples 1001.
int default_nsamples(long long npix) {
  const long long target = 1LL << 28;
  return static_cast<int>(std::max(1LL, std::min(256LL, target / std::max(npix, 1LL))));
}

void copy_simd_stats(xnh_stats& o, const unsigned long long* st) {
  o.interior_iters = st[kStIntIters];
  o.leaf_iters = st[kStLeafIters];
  o.leaf_steps = st[kStLeafSteps];
  o.round_iters = st[kStRoundIters];
  o.round_lanes = st[kStRoundLanes];
  o.shadow_interior_iters = st[kStShIntIters];
  o.shadow_leaf_iters = st[kStShLeafIters];
  o.shadow_leaf_steps = st[kStShLeafSteps];
  o.shadow_round_iters = st[kStShRoundIters];
  o.shadow_round_lanes = st[kStShRoundLanes];
}

int fail(xnh_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_OR_FAIL(ctx, expr)                                                              \
  do {                                                                                      \
    const hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                                   \
      return fail((ctx), XNH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

// hipMalloc that reports a refusal as XNH_E_ALLOC and clears the runtime's
// last-error state it leaves, which the next launch check (hipGetLastError)
// would otherwise report as a failure of that launch.
int raw_alloc(xnh_ctx* c, void** p, size_t bytes, const char* what) {
  *p = nullptr;
  const hipError_t e = hipMalloc(p, bytes);
  if (e == hipSuccess) return XNH_OK;
  (void)hipGetLastError();
  *p = nullptr;
  return fail(c, XNH_E_ALLOC, std::string(what) + ": hipMalloc: " + hipGetErrorString(e));
}

template <class T>
int dev_alloc(xnh_ctx* c, std::vector<void*>& owner, T** out, size_t count) {
  *out = nullptr;
  if (count == 0) count = 1;
  void* p = nullptr;
  const int rc = raw_alloc(c, &p, count * sizeof(T), "device buffer");
  if (rc != XNH_OK) return rc;
  owner.push_back(p);
  *out = static_cast<T*>(p);
  return XNH_OK;
}

template <class T>
int dev_upload(xnh_ctx* c, std::vector<void*>& owner, T** out, const std::vector<T>& host) {
  int rc = dev_alloc(c, owner, out, host.size());
  if (rc != XNH_OK) return rc;
  if (!host.empty()) {
    const hipError_t e = hipMemcpy(*out, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(c, XNH_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  return XNH_OK;
}

void free_all(std::vector<void*>& v) {
  for (void* p : v) hipFree(p);
  v.clear();
}

float4 f4(float x, float y, float z, float w) {
  float4 r;
  r.x = x; r.y = y; r.z = z; r.w = w;
  return r;
}
float bits(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
float ibits(int32_t i) {
  float f;
  std::memcpy(&f, &i, 4);
  return f;
}
uint32_t as_uint(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// Geometric normal of a light triangle, mesh.h:64-68 (host IEEE order).
void face_normal(const float p0[3], const float p1[3], const float p2[3], float out[3]) {
  const float a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
  const float b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
  const float c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const float len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  const float r = 1.f / len;
  out[0] = c[0] * r;
  out[1] = c[1] * r;
  out[2] = c[2] * r;
}

int flatten_upload(xnh_ctx* c, const xnh_flat_scene* fs, xnh_gpu_scene* g) {
  if (!c || !fs || !g) return XNH_E_INVALID;
  if (fs->num_meshes < 0 || fs->num_instances < 0 || fs->num_materials < 0 || fs->num_lights < 0 ||
      fs->num_dir_lights < 0 || fs->num_tlas_nodes < 0)
    return fail(c, XNH_E_INVALID, "negative count in flat scene");
  const int nm = fs->num_meshes;
  std::vector<uint32_t> node_base(nm), slot_base(nm), tri_base(nm), vert_base(nm);
  size_t nn = 0, ns = 0, nv = 0;
  for (int m = 0; m < nm; ++m) {
    const xnh_mesh& me = fs->meshes[m];
    if (me.num_triangles < 0 || me.num_vertices < 0 || me.num_nodes < 0)
      return fail(c, XNH_E_INVALID, "negative mesh count");
    if (me.num_triangles > 0 && me.num_nodes < 1) return fail(c, XNH_E_INVALID, "mesh without BLAS nodes");
    if ((nn & 1u) == 0) ++nn;   // odd root: child pairs on 64-byte boundaries
    node_base[m] = static_cast<uint32_t>(nn);
    slot_base[m] = static_cast<uint32_t>(ns);
    tri_base[m] = static_cast<uint32_t>(ns);
    vert_base[m] = static_cast<uint32_t>(nv);
    nn += static_cast<size_t>(me.num_nodes);
    ns += static_cast<size_t>(me.num_triangles);
    nv += static_cast<size_t>(me.num_vertices);
  }
  if (ns >= (size_t(1) << kEntShift) || nn + 2 >= (size_t(1) << kEntShift))
    return fail(c, XNH_E_INVALID, "too many triangles or BLAS nodes (limit 2^28)");
  // analytic primitives and their PrimBLAS
  const int np = fs->prims ? fs->num_prims : 0, npn = fs->prim_nodes ? fs->num_prim_nodes : 0;
  // (a primitive index below 2^31 never reaches the kPrimHit flag bit)
  if (np < 0 || npn < 0 || (np > 0 && (npn < 1 || !fs->prim_order)))
    return fail(c, XNH_E_INVALID, "bad primitive arrays");
  std::vector<float4> prims(3 * static_cast<size_t>(np)), pnodes(2 * static_cast<size_t>(npn));
  std::vector<uint32_t> porder(static_cast<size_t>(np));
  for (int i = 0; i < np; ++i) {
    const xnh_prim& pr = fs->prims[i];
    if (pr.type < XNH_PRIM_SPHERE || pr.type > XNH_PRIM_QUAD) return fail(c, XNH_E_INVALID, "unknown primitive type");
    if (pr.mat_id < 0 || pr.mat_id >= fs->num_materials) return fail(c, XNH_E_INVALID, "primitive material out of range");
    const float* q = pr.p;
    prims[3 * i] = f4(ibits(pr.type), ibits(pr.mat_id), q[0], q[1]);
    prims[3 * i + 1] = f4(q[2], q[3], q[4], q[5]);
    prims[3 * i + 2] = f4(q[6], q[7], q[8], q[9]);
    porder[i] = fs->prim_order[i];
    if (porder[i] >= static_cast<uint32_t>(np)) return fail(c, XNH_E_INVALID, "primitive index out of range");
  }
  for (int j = 0; j < npn; ++j) {
    const xnh_bvh_node& n = fs->prim_nodes[j];
    if (n.count > 0 && static_cast<uint64_t>(n.child) + n.count > static_cast<uint64_t>(np))
      return fail(c, XNH_E_INVALID, "PrimBLAS leaf outside primitive range");
    if (n.count == 0 && (n.child + 1 >= static_cast<uint32_t>(npn) || n.child <= static_cast<uint32_t>(j)))
      return fail(c, XNH_E_INVALID, "PrimBLAS child outside node range");
    pnodes[2 * j] = f4(n.lo[0], n.lo[1], n.lo[2], n.hi[0]);
    pnodes[2 * j + 1] = f4(n.hi[1], n.hi[2], bits(n.child), bits(n.count));
  }
  std::vector<float4> bnodes(2 * nn), tris(3 * ns), vnrm(nv);
  std::vector<uint32_t> bigcnt(std::max<size_t>(ns, 1), 0u);   // triangles of a large leaf, at its first slot
  std::vector<int32_t> vidx(3 * ns), tmat(ns);
  std::vector<bool> flat(nm, false);
  for (int m = 0; m < nm; ++m) {
    const xnh_mesh& me = fs->meshes[m];
    for (int j = 0; j < me.num_nodes; ++j) {
      const xnh_bvh_node& n = me.nodes[j];
      const uint32_t child = n.count > 0 ? slot_base[m] + n.child : node_base[m] + n.child;
      if (n.count > 0 && n.child + n.count > static_cast<uint32_t>(me.num_triangles))
        return fail(c, XNH_E_INVALID, "BLAS leaf outside prim range");
      // children after their parent (the builders' pre-order allocation):
      // every walk then moves to strictly larger indices and must end
      if (n.count == 0 && (n.child + 1 >= static_cast<uint32_t>(me.num_nodes) || n.child <= static_cast<uint32_t>(j)))
        return fail(c, XNH_E_INVALID, "BLAS child outside node range");
      // the node's stack entry, precomputed (see dtrace.h): interior: its
      // left child; leaf: count << kEntShift | first slot, a large leaf
      // with the count kept in bigcnt at that slot
      const uint32_t entry = n.count == 0 ? child : (std::min(n.count, kEntBigLeaf) << kEntShift) | child;
      if (n.count >= kEntBigLeaf) bigcnt[child] = n.count;
      bnodes[2 * (node_base[m] + j)] = f4(n.lo[0], n.lo[1], n.lo[2], n.hi[0]);
      bnodes[2 * (node_base[m] + j) + 1] = f4(n.hi[1], n.hi[2], bits(child), bits(entry));
    }
    for (int s = 0; s < me.num_triangles; ++s) {
      const uint32_t t = me.prims[s];
      if (t >= static_cast<uint32_t>(me.num_triangles)) return fail(c, XNH_E_INVALID, "prim index out of range");
      int iv[3];
      for (int k = 0; k < 3; ++k) {
        iv[k] = me.indices[3 * t + k];
        if (iv[k] < 0 || iv[k] >= me.num_vertices) return fail(c, XNH_E_INVALID, "vertex index out of range");
      }
      const float v0[3] = {me.vx[iv[0]], me.vy[iv[0]], me.vz[iv[0]]};
      const float v1[3] = {me.vx[iv[1]], me.vy[iv[1]], me.vz[iv[1]]};
      const float v2[3] = {me.vx[iv[2]], me.vy[iv[2]], me.vz[iv[2]]};
      const float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
      const float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
      const size_t g = slot_base[m] + s;
      tris[3 * g] = f4(v0[0], v0[1], v0[2], e1[0]);
      tris[3 * g + 1] = f4(e1[1], e1[2], e2[0], e2[1]);
      tris[3 * g + 2] = f4(e2[2], bits(t), 0.f, 0.f);
    }
    for (int t = 0; t < me.num_triangles; ++t) {
      for (int k = 0; k < 3; ++k) vidx[3 * (tri_base[m] + t) + k] = me.indices[3 * t + k] + static_cast<int32_t>(vert_base[m]);
      const int mid = me.mat_ids[t];
      if (mid < 0 || mid >= fs->num_materials) return fail(c, XNH_E_INVALID, "material id out of range");
      tmat[tri_base[m] + t] = mid;
    }
    flat[m] = !(me.nx && me.ny && me.nz);
    for (int v = 0; v < me.num_vertices; ++v)
      vnrm[vert_base[m] + v] = flat[m] ? f4(0.f, 0.f, 0.f, 0.f) : f4(me.nx[v], me.ny[v], me.nz[v], 0.f);
  }
  // shading class of each mesh: its materials' common class (see ShadeClass)
  auto mat_class = [&](int mid) -> uint32_t {
    const xnh_material& m = fs->materials[mid];
    if (m.flags & XNH_MAT_DELTA) return kShadeDelta;
    if (m.flags & XNH_MAT_SUBSURFACE) return kShadeSubsurface;
    if (m.flags & XNH_MAT_TRANSMISSIVE) return kShadeRoughGlass;
    if (m.flags & XNH_MAT_CONDUCTOR) return kShadeConductor;
    return m.f0[0] == 0.f ? kShadeLambert : kShadeDielectric;
  };
  std::vector<uint32_t> mesh_class(nm, kShadeMixed);
  for (int m = 0; m < nm; ++m) {
    const xnh_mesh& me = fs->meshes[m];
    uint32_t k = me.num_triangles > 0 ? mat_class(me.mat_ids[0]) : kShadeMixed;
    for (int t = 1; t < me.num_triangles && k != kShadeMixed; ++t)
      if (mat_class(me.mat_ids[t]) != k) k = kShadeMixed;
    mesh_class[m] = k;
  }
  // TLAS nodes with their child-order key
  const int ntn = fs->num_tlas_nodes;
  std::vector<float4> tnodes(3 * static_cast<size_t>(ntn));
  for (int j = 0; j < ntn; ++j) {
    const xnh_bvh_node& n = fs->tlas_nodes[j];
    tnodes[3 * j] = f4(n.lo[0], n.lo[1], n.lo[2], n.hi[0]);
    tnodes[3 * j + 1] = f4(n.hi[1], n.hi[2], bits(n.child), bits(n.count));
    float4 key = f4(0.f, 0.f, bits(0u), 0.f);
    if (n.count == 0) {
      if (n.child + 1 >= static_cast<uint32_t>(ntn) || n.child <= static_cast<uint32_t>(j))
        return fail(c, XNH_E_INVALID, "TLAS child out of range");
      const float ex = n.hi[0] - n.lo[0], ey = n.hi[1] - n.lo[1], ez = n.hi[2] - n.lo[2];
      const int ax = (ex > ey && ex > ez) ? 0 : (ey > ez ? 1 : 2);
      const xnh_bvh_node& l = fs->tlas_nodes[n.child];
      const xnh_bvh_node& r = fs->tlas_nodes[n.child + 1];
      key = f4((l.lo[ax] + l.hi[ax]) * 0.5f, (r.lo[ax] + r.hi[ax]) * 0.5f, bits(static_cast<uint32_t>(ax)), 0.f);
    } else if (n.child + n.count > static_cast<uint32_t>(fs->num_instances)) {
      return fail(c, XNH_E_INVALID, "TLAS leaf outside instance range");
    }
    tnodes[3 * j + 2] = key;
  }
  std::vector<uint32_t> torder(fs->tlas_inst_indices, fs->tlas_inst_indices + fs->num_instances);
  // per-instance light lookup tables: first light naming (instance id, triangle)
  std::vector<int32_t> lmap;
  std::vector<float4> inst(4 * static_cast<size_t>(fs->num_instances));
  std::vector<uint32_t> ikind(std::max(fs->num_instances, 1), kShadeMixed);
  for (int i = 0; i < fs->num_instances; ++i) {
    const xnh_instance& in = fs->instances[i];
    if (torder.size() > 0 && torder[i] >= static_cast<uint32_t>(fs->num_instances))
      return fail(c, XNH_E_INVALID, "TLAS instance index out of range");
    const float* w = in.world_to_local;
    if (in.mesh == XNH_PRIM_GEOMETRY) {   // the primitives (no area lights on them)
      if (npn < 1) return fail(c, XNH_E_INVALID, "primitive instance without primitives");
      inst[4 * i] = f4(w[0], w[1], w[2], w[3]);
      inst[4 * i + 1] = f4(w[4], w[5], w[6], w[7]);
      inst[4 * i + 2] = f4(w[8], w[9], w[10], w[11]);
      inst[4 * i + 3] = f4(bits(kPrimRoot), ibits(in.instance_id), bits(0u), ibits(-1));
      continue;
    }
    if (in.mesh < 0 || in.mesh >= nm) return fail(c, XNH_E_INVALID, "instance mesh out of range");
    ikind[i] = mesh_class[in.mesh];
    int32_t lm = -1;
    const int ntri = fs->meshes[in.mesh].num_triangles;
    for (int l = 0; l < fs->num_lights; ++l) {
      if (static_cast<int64_t>(fs->lights[l].mesh_id) != static_cast<int64_t>(in.instance_id)) continue;
      if (lm < 0) {
        lm = static_cast<int32_t>(lmap.size());
        lmap.resize(lmap.size() + std::max(ntri, 1), -1);
      }
      const uint32_t t = fs->lights[l].tri_idx;
      if (t < static_cast<uint32_t>(ntri) && lmap[lm + t] < 0) lmap[lm + t] = l;
    }
    uint32_t root = kNoRoot;
    if (fs->meshes[in.mesh].num_nodes > 0) {
      root = as_uint(bnodes[2 * node_base[in.mesh] + 1].w);   // the root node's entry
    }
    inst[4 * i] = f4(w[0], w[1], w[2], w[3]);
    inst[4 * i + 1] = f4(w[4], w[5], w[6], w[7]);
    inst[4 * i + 2] = f4(w[8], w[9], w[10], w[11]);
    inst[4 * i + 3] = f4(bits(root), ibits(in.instance_id), bits(tri_base[in.mesh] | (flat[in.mesh] ? kFlatBit : 0u)),
                         ibits(lm));
  }
  // lights: corners in the coordinates of mesh `mesh_id` (cpu_tiled.cpp:93-106)
  std::vector<float4> lights(4 * static_cast<size_t>(fs->num_lights));
  for (int l = 0; l < fs->num_lights; ++l) {
    const xnh_light& L = fs->lights[l];
    if (L.mesh_id >= static_cast<uint32_t>(nm)) return fail(c, XNH_E_INVALID, "light mesh out of range");
    const xnh_mesh& me = fs->meshes[L.mesh_id];
    if (L.tri_idx >= static_cast<uint32_t>(me.num_triangles)) return fail(c, XNH_E_INVALID, "light triangle out of range");
    float p[3][3];
    for (int k = 0; k < 3; ++k) {
      const int v = me.indices[3 * L.tri_idx + k];
      p[k][0] = me.vx[v];
      p[k][1] = me.vy[v];
      p[k][2] = me.vz[v];
    }
    float n[3];
    face_normal(p[0], p[1], p[2], n);
    lights[4 * l] = f4(p[0][0], p[0][1], p[0][2], L.area);
    lights[4 * l + 1] = f4(p[1][0], p[1][1], p[1][2], L.emission[0]);
    lights[4 * l + 2] = f4(p[2][0], p[2][1], p[2][2], L.emission[1]);
    lights[4 * l + 3] = f4(n[0], n[1], n[2], L.emission[2]);
  }
  std::vector<float4> mats(kMatVecs * static_cast<size_t>(fs->num_materials));
  for (int i = 0; i < fs->num_materials; ++i) {
    const xnh_material& m = fs->materials[i];
    float4* r = mats.data() + kMatVecs * i;
    r[0] = f4(m.base_color[0], m.base_color[1], m.base_color[2], m.metallic);
    r[1] = f4(m.ior, m.alpha, m.alpha_x, m.alpha_y);
    r[2] = f4(m.f0[0], m.f0[1], m.f0[2], m.subsurface);
    r[3] = f4(m.subsurface_color[0], m.subsurface_color[1], m.subsurface_color[2], bits(m.flags));
    // the material-only parts of the multiple-scattering compensation
    // (bsdf.h:212-249), in the device's float operation order; the two
    // transcendentals evaluated in double and rounded once, as the device
    // does for its own (device/dmath.h cos_r etc.)
    const float a = m.alpha;
    const float k1 = 5.0f * static_cast<float>(std::exp(static_cast<double>(-2.69f * a)));
    const float k2 = 1.0f + 22.7f * static_cast<float>(std::pow(static_cast<double>(a), 1.5));
    const float ea = 1.0f / (1.0f + 0.3037f * a + 0.1218f * a * a);
    float fms[3];
    const float fav0 = (1.0f + m.f0[0] * 20.f) * (1.f / 21.f);
    float den = 1.f - fav0 * (1.f - ea);
    den = (den < 1e-4f) ? 1e-4f : den;
    for (int k = 0; k < 3; ++k) {
      const float fav = (1.0f + m.f0[k] * 20.f) * (1.f / 21.f);
      fms[k] = fav * fav * ea / den;
    }
    r[4] = f4(k1, k2, fms[0], fms[1]);
    r[5] = f4(fms[2], 0.f, 0.f, 0.f);
  }
  std::vector<float4> suns(2 * static_cast<size_t>(fs->num_dir_lights));
  for (int i = 0; i < fs->num_dir_lights; ++i) {
    const xnh_dir_light& d = fs->dir_lights[i];
    suns[2 * i] = f4(d.direction[0], d.direction[1], d.direction[2], 0.f);
    suns[2 * i + 1] = f4(d.emission[0], d.emission[1], d.emission[2], 0.f);
  }
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  free_all(g->allocs);
  DevScene S{};
  float4 *p_bn, *p_tr, *p_tn, *p_in, *p_vn, *p_li, *p_ma, *p_su, *p_pr = nullptr, *p_pn = nullptr;
  uint32_t *p_to, *p_po = nullptr, *p_bc, *p_ik;
  int32_t *p_vi, *p_tm, *p_lm;
  int rc = dev_upload(c, g->allocs, &p_bn, bnodes);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_bc, bigcnt);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_tr, tris);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_tn, tnodes);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_to, torder);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_in, inst);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_vi, vidx);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_vn, vnrm);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_tm, tmat);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_lm, lmap);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_ik, ikind);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_li, lights);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_ma, mats);
  if (rc == XNH_OK) rc = dev_upload(c, g->allocs, &p_su, suns);
  if (rc == XNH_OK && np > 0) rc = dev_upload(c, g->allocs, &p_pr, prims);
  if (rc == XNH_OK && np > 0) rc = dev_upload(c, g->allocs, &p_pn, pnodes);
  if (rc == XNH_OK && np > 0) rc = dev_upload(c, g->allocs, &p_po, porder);
  if (rc != XNH_OK) {
    free_all(g->allocs);
    return rc;
  }
  S.bnodes = p_bn;
  S.bigcnt = p_bc;
  S.tris = p_tr;
  S.tnodes = p_tn;
  S.torder = p_to;
  S.inst = p_in;
  S.vidx = p_vi;
  S.vnrm = p_vn;
  S.tmat = p_tm;
  S.lmap = p_lm;
  S.lights = p_li;
  S.mats = p_ma;
  S.suns = p_su;
  S.prims = p_pr;
  S.pnodes = p_pn;
  S.porder = p_po;
  S.ikind = p_ik;
  S.num_pnodes = np > 0 ? npn : 0;
  S.num_tnodes = ntn;
  S.num_lights = fs->num_lights;
  S.num_suns = fs->num_dir_lights;
  for (int k = 0; k < 3; ++k) {
    S.sky.zenith[k] = fs->sky.zenith[k];
    S.sky.horizon[k] = fs->sky.horizon[k];
    S.sky.ground[k] = fs->sky.ground[k];
  }
  S.sky.sharpness = fs->sky.sharpness;
  S.sky.intensity = fs->sky.intensity;
  g->S = S;
  g->bytes = 0;
  for (const auto* v : {&bnodes, &tris, &tnodes, &inst, &vnrm, &lights, &mats, &suns, &prims, &pnodes})
    g->bytes += v->size() * 16;
  g->bytes += (torder.size() + vidx.size() + tmat.size() + lmap.size() + porder.size() + bigcnt.size() +
               ikind.size()) * 4;
  return XNH_OK;
}

// Queues for the current tuning (block, lds_depth, nsamples, grid_trace).
// The previous queues are freed first: at the default batch one set is a
// quarter of the device, so two sets need not fit together. On failure
// nothing is left allocated and B.capacity is 0.
int alloc_waves(xnh_ctx* c) {
  free_all(c->wave_allocs);
  const int npix = c->cfg.width * c->cfg.height;
  const size_t cap = static_cast<size_t>(npix) * c->nsamples;
  // queue entries: the slots plus the per-chunk segment slack (dev_types.h)
  const size_t qcap = cap + static_cast<size_t>(kQueueSlack);
  WaveBuffers& B = c->B;
  B = WaveBuffers{};
  B.capacity = static_cast<int32_t>(cap);
  B.qcap = static_cast<int32_t>(qcap);
  int rc = XNH_OK;
  for (int k = 0; k < 2 && rc == XNH_OK; ++k) {
    rc = dev_alloc(c, c->wave_allocs, &B.q[k].ray0, qcap);
    if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.q[k].ray1, qcap);
    if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.q[k].thr, qcap);
    if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.q[k].st, qcap);
  }
  if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.hit, qcap);
  if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.hit_slot, qcap);
  if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.sh0, 2 * qcap);
  if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.sh1, 2 * qcap);
  if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.sh_c, 2 * qcap);
  if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.l_term, cap);
  if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.l_nee, 2 * cap);
  if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.counters, static_cast<size_t>(kNumCtr));
  if (rc == XNH_OK) rc = dev_alloc(c, c->wave_allocs, &B.stats, static_cast<size_t>(kNumStat));
  if (rc == XNH_OK) {
    const size_t spill = static_cast<size_t>(kMaxStack - c->lds_depth) * c->grid_trace * c->block;
    rc = dev_alloc(c, c->wave_allocs, &B.spill, spill);
  }
  if (rc != XNH_OK) {
    free_all(c->wave_allocs);
    B = WaveBuffers{};
    return rc;
  }
  HIP_OR_FAIL(c, hipMemset(B.counters, 0, sizeof(uint32_t) * kNumCtr));
  HIP_OR_FAIL(c, hipMemset(B.stats, 0, sizeof(unsigned long long) * kNumStat));
  return XNH_OK;
}

// The running per-pixel sums and the averaging buffer, allocated once.
int alloc_pixels(xnh_ctx* c) {
  const size_t npix = static_cast<size_t>(c->cfg.width) * c->cfg.height;
  int rc = dev_alloc(c, c->pix_allocs, &c->accum, npix);
  if (rc == XNH_OK) rc = dev_alloc(c, c->pix_allocs, &c->d_out, npix * 3);
  if (rc != XNH_OK) return rc;
  HIP_OR_FAIL(c, hipMemset(c->accum, 0, sizeof(float4) * npix));
  return XNH_OK;
}

// Grows the job / chunk tables; a reallocation invalidates the cached graph
// (its kernel arguments hold the table addresses).
int ensure_tables(xnh_ctx* c, size_t njobs, size_t nbatches) {
  if (njobs > c->jobs_cap) {
    if (c->d_jobs) hipFree(c->d_jobs);
    c->d_jobs = nullptr;
    c->jobs_cap = 0;
    c->table_gen++;
    size_t cap = std::max<size_t>(njobs, 64);
    void* p = nullptr;
    const int rc = raw_alloc(c, &p, cap * sizeof(DevJob), "job table");
    if (rc != XNH_OK) return rc;
    c->d_jobs = static_cast<DevJob*>(p);
    c->jobs_cap = cap;
    c->table_gen++;
  }
  if (nbatches > c->batches_cap) {
    if (c->d_chunks) hipFree(c->d_chunks);
    if (c->d_nchunks) hipFree(c->d_nchunks);
    c->d_chunks = nullptr;
    c->d_nchunks = nullptr;
    c->batches_cap = 0;
    c->table_gen++;
    size_t cap = std::max<size_t>(nbatches, 256);
    void *p = nullptr, *q = nullptr;
    int rc = raw_alloc(c, &p, cap * kMaxChunks * sizeof(DevChunk), "chunk table");
    if (rc != XNH_OK) return rc;
    c->d_chunks = static_cast<DevChunk*>(p);
    rc = raw_alloc(c, &q, cap * sizeof(int32_t), "chunk counts");
    if (rc != XNH_OK) return rc;
    c->d_nchunks = static_cast<int32_t*>(q);
    c->batches_cap = cap;
    c->table_gen++;
  }
  c->B.jobs = c->d_jobs;
  c->B.chunks = c->d_chunks;
  c->B.nchunks = c->d_nchunks;
  return XNH_OK;
}

// Enqueues one batch: the whole bounce loop for `depth` bounces.
int enqueue_batch(xnh_ctx* c, int depth, size_t* ev) {
  hipStream_t s = c->stream;
  const int cap = c->B.capacity;
  HIP_OR_FAIL(c, launch_raygen(s, c->B, cap));
  for (int d = 0; d < depth; ++d) {
    const int qi = d & 1;
    if (ev) HIP_OR_FAIL(c, hipEventRecord(c->prof_events[(*ev)++], s));
    HIP_OR_FAIL(c, launch_closest(s, c->B, qi, 2 * d, c->grid_trace, c->block, c->lds_depth, c->counting, c->prims,
                                  d == 0));
    if (ev) HIP_OR_FAIL(c, hipEventRecord(c->prof_events[(*ev)++], s));
    if (c->guide.on && d < c->B.guide_depth)
      HIP_OR_FAIL(c, launch_guide_infer(s, c->B, qi, d, c->guide.net, c->guide.feat, c->grid_shade * 2));
    if (ev) HIP_OR_FAIL(c, hipEventRecord(c->prof_events[(*ev)++], s));
    HIP_OR_FAIL(c, launch_shade(s, c->B, qi, d, c->grid_shade));
    if (ev) HIP_OR_FAIL(c, hipEventRecord(c->prof_events[(*ev)++], s));
    if (ev) HIP_OR_FAIL(c, hipEventRecord(c->prof_events[(*ev)++], s));
    HIP_OR_FAIL(c, launch_shadow(s, c->B, 2 * d + 1, c->grid_trace, c->block, c->lds_depth, c->counting, c->prims));
    if (ev) HIP_OR_FAIL(c, hipEventRecord(c->prof_events[(*ev)++], s));
  }
  HIP_OR_FAIL(c, launch_accumulate(s, c->B, cap));
  HIP_OR_FAIL(c, launch_advance(s, c->B.counters));
  return XNH_OK;
}

int ensure_graph(xnh_ctx* c, int depth) {
  if (c->graph && c->graph_depth == depth && c->graph_gen == c->table_gen) return XNH_OK;
  if (c->graph) {
    hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  HIP_OR_FAIL(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_batch(c, depth, nullptr);
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc != XNH_OK) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(c, XNH_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  const hipError_t e2 = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e2 != hipSuccess) {
    c->graph = nullptr;
    return fail(c, XNH_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
  }
  c->graph_depth = depth;
  c->graph_gen = c->table_gen;
  return XNH_OK;
}

// Room for the per-work-group weight gradients of a training launch of
// `blocks` work-groups.
static int ensure_wpart(xnh_ctx* c, int blocks) {
  GuideState& g = c->guide;
  if (blocks <= g.w_part_blocks) return XNH_OK;
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  if (g.w_part) {
    for (auto& p : g.allocs)
      if (p == g.w_part) p = nullptr;
    HIP_OR_FAIL(c, hipFree(g.w_part));
    g.w_part = nullptr;
    g.w_part_blocks = 0;
  }
  const int rc = dev_alloc(c, g.allocs, &g.w_part, static_cast<size_t>(blocks) * kWTotal);
  if (rc != XNH_OK) return rc;
  g.w_part_blocks = blocks;
  return XNH_OK;
}

// Guiding steps on the batch just rendered: records -> training batch (all
// counts stay on the device) -> `train_steps` fused forward/backward + Adam
// steps on equal slices of it; then the guide's MIS probability for the
// next batch (ramped over the warm-up steps).
int guide_train_batch(xnh_ctx* c) {
  GuideState& g = c->guide;
  hipStream_t s = c->stream;
  const int R = g.cfg.train_batch;
  const int K = g.cfg.train_steps;
  HIP_OR_FAIL(c, hipMemsetAsync(g.t_lam_sum, 0, sizeof(float), s));
  HIP_OR_FAIL(c, hipMemsetAsync(g.lam_sum_all, 0, sizeof(float), s));
  HIP_OR_FAIL(c, hipMemsetAsync(g.t_n, 0, sizeof(uint32_t), s));
  // expected count 0.9 R, so that the cap seldom clips the draw
  HIP_OR_FAIL(c, launch_guide_collect(s, c->B, g.lam_all, g.lam_sum_all, 0.9f * static_cast<float>(R), g.collects++,
                                      g.t_pos, g.t_nrm, g.t_dir, g.t_lam, g.t_lam_sum, g.t_n,
                                      static_cast<uint32_t>(R)));
  const float b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
  const float lr = g.cfg.learning_rate;
  {
    const int rc = ensure_wpart(c, guide_train_blocks(R, true, K));
    if (rc != XNH_OK) return rc;
  }
  for (int k = 0; k < K; ++k) {
    HIP_OR_FAIL(c, hipMemsetAsync(g.loss, 0, sizeof(float), s));
    HIP_OR_FAIL(c, launch_guide_train(s, g.net, R, g.t_pos, g.t_nrm, g.t_dir, g.t_lam, 1.f, g.t_n, g.t_lam_sum, k, K,
                                      g.grid_g, g.w_part, g.w_g, g.loss));
    g.step++;
    const float bc1 = 1.f - static_cast<float>(std::pow(double(b1), g.step));
    const float bc2 = 1.f - static_cast<float>(std::pow(double(b2), g.step));
    HIP_OR_FAIL(c, launch_guide_adam(s, g.grid_m, g.grid_h, g.grid_g, g.grid_am, g.grid_av, g.ngrid, lr, b1, b2, eps,
                                     bc1, bc2, 1.f, g.t_n, static_cast<uint32_t>(R), K));
    HIP_OR_FAIL(c, launch_guide_adam(s, g.w_m, g.w_h, g.w_g, g.w_am, g.w_av, size_t(kWTotal), lr, b1, b2, eps, bc1,
                                     bc2, 1.f, g.t_n, static_cast<uint32_t>(R), K));
    HIP_OR_FAIL(c, launch_tile_weights(s, g.w_h, g.w_t));
  }
  const float a = g.cfg.warmup_steps > 0
                      ? g.cfg.guide_fraction * std::min(1.f, static_cast<float>(g.step) / g.cfg.warmup_steps)
                      : g.cfg.guide_fraction;
  if (a != g.alpha_host) {
    g.alpha_host = a;
    HIP_OR_FAIL(c, hipMemcpyAsync(g.alpha_dev, &g.alpha_host, sizeof(float), hipMemcpyHostToDevice, s));
  }
  return XNH_OK;
}

struct JobSpec {
  const DevScene* S;
  xnh_camera cam;
  int W, H, spp_first, spp_count, maxb, minb;
  float4* accum;
};

// Renders a list of jobs; sums are added to each job's accumulator.
// Paths per batch while guiding: the guide learns between batches, and the
// caustic room renders best at equal time with batches of about this size
// (1280x720: 72 samples, about 40 updates in 8 s) and one training step per
// 2^22 of their paths; the full capacity trained too seldom.
constexpr size_t kGuidedBatchPaths = size_t(1) << 26;

// Queue entries a guided batch uses at most: kGuidedBatchPaths, or one
// frame at the context's resolution if that is larger, within the queue
// capacity. The guide's per-entry buffers are sized by it.
size_t guided_cap(const xnh_ctx* c) {
  const size_t npix = static_cast<size_t>(c->cfg.width) * c->cfg.height;
  return std::min(static_cast<size_t>(c->B.capacity), std::max(kGuidedBatchPaths, npix));
}

// The guide's fields of the wavefront buffers (after the queues are
// (re)allocated or guiding is switched on or off).
void attach_guide(xnh_ctx* c) {
  const GuideState& g = c->guide;
  WaveBuffers& B = c->B;
  B.mix = g.on ? g.mix : nullptr;
  B.mix_stride = g.on ? g.mix_stride : 0u;
  B.rec = g.on ? g.rec : nullptr;
  B.rec_cap = g.on ? g.rec_cap : 0u;
  B.rec_every = g.on ? g.rec_every : 0u;
  B.guide_alpha = g.on ? g.cfg.guide_fraction : 0.f;
  B.alpha_now = g.on ? g.alpha_dev : nullptr;
  B.guide_depth = g.on ? g.cfg.guide_depth : 0;
}

int render_jobs(xnh_ctx* c, const std::vector<JobSpec>& jobs) {
  size_t cap = static_cast<size_t>(c->B.capacity);
  if (cap == 0)
    return fail(c, XNH_E_ALLOC, "no wavefront queues: a re-tuning failed and the previous tuning could not be "
                                "restored (a later xnh_set_tuning that succeeds makes the context usable again)");
  // queue entries a batch may use (its chunks' segments)
  size_t qlimit = static_cast<size_t>(c->B.qcap);
  if (c->guide.on) {
    cap = c->guide.slot_cap;
    qlimit = c->guide.mix_stride;
  }
  // batch schedule: greedy, one chunk per job per batch
  std::vector<DevChunk> chunks;
  std::vector<int32_t> counts;
  size_t used = 0, qused = 0;
  int in_batch = 0;
  auto new_batch = [&]() {
    counts.push_back(0);
    chunks.resize(counts.size() * kMaxChunks);
    used = 0;
    qused = 0;
    in_batch = 0;
  };
  int depth = 0;
  for (size_t j = 0; j < jobs.size(); ++j) {
    const JobSpec& js = jobs[j];
    const size_t npix = static_cast<size_t>(js.W) * js.H;
    if (npix > cap)
      return fail(c, XNH_E_INVALID, c->guide.on ? "job larger than a guided batch (the larger of 2^26 paths and one "
                                                  "frame at the context's resolution)"
                                                : "job larger than the path capacity");
    depth = std::max(depth, js.maxb);
    int remaining = js.spp_count, next = js.spp_first;
    while (remaining > 0) {
      if (counts.empty() || in_batch == kMaxChunks || used + npix > cap) new_batch();
      const int ns = static_cast<int>(std::min<size_t>(remaining, (cap - used) / npix));
      DevChunk& ck = chunks[(counts.size() - 1) * kMaxChunks + in_batch];
      ck = DevChunk{};
      ck.job = static_cast<int32_t>(j);
      ck.spp_first = next;
      ck.nsamples = ns;
      ck.npix = static_cast<int32_t>(npix);
      ck.base = static_cast<int32_t>(used);
      ck.cap = static_cast<int32_t>(npix * ns);
      ck.qseg = static_cast<int32_t>(qseg_for(ck.cap));
      ck.qbase = static_cast<int32_t>(qused);
      // Depth-0 order (camera_path): scenes of many instances walk their
      // TLAS and several BLAS per camera ray and gain from the one-pixel
      // footprint of 64 samples per block (instancing, 4096 instances:
      // +8.7 % over 8 x 8); in smaller scenes the depth-0 per-slot writes,
      // then 64 separate lines per wave, cost more than the coherence gains
      // (36-triangle Cornell box -1.3 %; a 1.3M-triangle two-instance scene
      // unchanged), and 8 samples of 8 pixels are taken. Guided batches keep
      // slot order (a pixel's first sample is then entry qbase + t, where
      // its mixture is stored).
      ck.cam_lg = 0;
      if (!c->guide.on) {
        const int lg = js.S->num_tnodes >= 127 ? 6 : 3;   // 127 TLAS nodes: 64 instances
        if (ns % (1 << lg) == 0 && npix % (size_t(1) << (6 - lg)) == 0) ck.cam_lg = lg;
        else if (ns % 8 == 0 && npix % 8 == 0) ck.cam_lg = 3;
      }
      ck.sdiv = fast_div(static_cast<uint32_t>(ck.cam_lg ? npix >> (6 - ck.cam_lg) : npix));
      qused += static_cast<size_t>(kQSeg) * static_cast<size_t>(ck.qseg);
      if (qused > qlimit) return fail(c, XNH_E_INVALID, "queue segments exceed the queue capacity");
      counts.back() = ++in_batch;
      used += npix * ns;
      remaining -= ns;
      next += ns;
    }
  }
  const size_t nb = counts.size();
  int rc = ensure_tables(c, std::max<size_t>(jobs.size(), 1), std::max<size_t>(nb, 1));
  if (rc != XNH_OK) return rc;
  std::vector<DevJob> dj(jobs.size());
  bool prims = false;
  for (size_t j = 0; j < jobs.size(); ++j) {
    DevJob& d = dj[j];
    d = DevJob{};
    d.S = *jobs[j].S;
    prims = prims || d.S.num_pnodes > 0;
    for (int k = 0; k < 3; ++k) {
      d.cam.eye[k] = jobs[j].cam.eye[k];
      d.cam.llc[k] = jobs[j].cam.lower_left[k];
      d.cam.horiz[k] = jobs[j].cam.horizontal[k];
      d.cam.vert[k] = jobs[j].cam.vertical[k];
    }
    d.W = jobs[j].W;
    d.band = fast_div(static_cast<uint32_t>(8 * jobs[j].W));
    d.H = jobs[j].H;
    d.minb = jobs[j].minb;
    d.maxb = jobs[j].maxb;
    d.accum = jobs[j].accum;
  }
  if (prims != c->prims) {
    c->prims = prims;
    c->table_gen++;   // the captured graph holds the other traversal variants
  }
  hipStream_t s = c->stream;
  HIP_OR_FAIL(c, hipStreamSynchronize(s));
  if (!dj.empty()) HIP_OR_FAIL(c, hipMemcpy(c->d_jobs, dj.data(), dj.size() * sizeof(DevJob), hipMemcpyHostToDevice));
  if (nb) {
    HIP_OR_FAIL(c, hipMemcpy(c->d_chunks, chunks.data(), chunks.size() * sizeof(DevChunk), hipMemcpyHostToDevice));
    HIP_OR_FAIL(c, hipMemcpy(c->d_nchunks, counts.data(), counts.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  HIP_OR_FAIL(c, hipMemsetAsync(c->B.stats, 0, sizeof(unsigned long long) * kNumStat, s));
  HIP_OR_FAIL(c, hipMemsetAsync(c->B.counters, 0, sizeof(uint32_t) * 16, s));   // batch index, overflow
  const size_t nev = c->profiling ? nb * depth * 6 : 0;
  while (c->prof_events.size() < nev) {
    hipEvent_t e;
    HIP_OR_FAIL(c, hipEventCreate(&e));
    c->prof_events.push_back(e);
  }
  const bool graph = c->use_graph && !c->profiling && nb > 0;
  if (graph) {
    rc = ensure_graph(c, depth);
    if (rc != XNH_OK) return rc;
  }
  HIP_OR_FAIL(c, hipEventRecord(c->ev_begin, s));
  size_t ev = 0;
  for (size_t b = 0; b < nb; ++b) {
    if (graph) {
      HIP_OR_FAIL(c, hipGraphLaunch(c->graph, s));
    } else {
      rc = enqueue_batch(c, depth, c->profiling ? &ev : nullptr);
      if (rc != XNH_OK) return rc;
    }
    if (c->guide.on && c->guide.train) {
      rc = guide_train_batch(c);
      if (rc != XNH_OK) return rc;
    }
  }
  HIP_OR_FAIL(c, hipEventRecord(c->ev_end, s));
  unsigned long long st[kNumStat];
  uint32_t ovf = 0;
  HIP_OR_FAIL(c, hipMemcpyAsync(st, c->B.stats, sizeof st, hipMemcpyDeviceToHost, s));
  HIP_OR_FAIL(c, hipMemcpyAsync(&ovf, c->B.counters + kOverflow, sizeof ovf, hipMemcpyDeviceToHost, s));
  uint32_t dog = 0, segovf = 0;
  HIP_OR_FAIL(c, hipMemcpyAsync(&dog, c->B.counters + kWatchdog, sizeof dog, hipMemcpyDeviceToHost, s));
  HIP_OR_FAIL(c, hipMemcpyAsync(&segovf, c->B.counters + kSegOverflow, sizeof segovf, hipMemcpyDeviceToHost, s));
  HIP_OR_FAIL(c, hipStreamSynchronize(s));
  xnh_stats& o = c->stats;
  o = xnh_stats{};
  o.paths = st[kStPaths];
  o.closest_rays = st[kStClosest];
  o.shadow_rays = st[kStShadow];
  o.node_tests = st[kStNodeTests];
  o.interior = st[kStInterior];
  o.tri_tests = st[kStTriTests];
  o.inst_visits = st[kStInstVisits];
  o.shadow_node_tests = st[kStShNodeTests];
  o.shadow_interior = st[kStShInterior];
  o.shadow_tri_tests = st[kStShTriTests];
  o.shadow_inst_visits = st[kStShInstVisits];
  o.nonfinite = st[kStNonFinite];
  copy_simd_stats(o, st);
  float ms = 0.f;
  HIP_OR_FAIL(c, hipEventElapsedTime(&ms, c->ev_begin, c->ev_end));
  o.render_ms = ms;
  o.closest_launches = static_cast<int32_t>(nb * depth);
  o.stack_overflows = static_cast<int32_t>(ovf);
  if (c->profiling) {
    for (size_t k = 0; k + 5 < ev; k += 6) {
      float a = 0.f, b = 0.f, d = 0.f;
      HIP_OR_FAIL(c, hipEventElapsedTime(&a, c->prof_events[k], c->prof_events[k + 1]));
      HIP_OR_FAIL(c, hipEventElapsedTime(&b, c->prof_events[k + 2], c->prof_events[k + 3]));
      HIP_OR_FAIL(c, hipEventElapsedTime(&d, c->prof_events[k + 4], c->prof_events[k + 5]));
      o.closest_ms += a;
      o.shade_ms += b;
      o.shadow_ms += d;
    }
  }
  if (dog) return fail(c, XNH_E_HIP, "traversal watchdog: " + std::to_string(dog) + " wave(s) reached the iteration bound");
  if (segovf)
    return fail(c, XNH_E_HIP, "queue segment overflow: " + std::to_string(segovf) + " append(s) dropped (internal error)");
  if (ovf) return fail(c, XNH_E_OVERFLOW, "traversal stack overflow");
  return XNH_OK;
}

}  // namespace

extern "C" {

int xnh_create(const xnh_config* cfg, xnh_ctx** out) {
  if (!cfg || !out) return XNH_E_INVALID;
  *out = nullptr;
  if (cfg->width <= 0 || cfg->height <= 0 || cfg->max_bounces < 0 || cfg->min_bounces < 0) return XNH_E_INVALID;
  if (cfg->max_bounces > 64) return XNH_E_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return XNH_E_NO_DEVICE;
  xnh_ctx* c = new (std::nothrow) xnh_ctx();
  if (!c) return XNH_E_ALLOC;
  c->cfg = *cfg;
  c->block = cfg->block_size > 0 ? cfg->block_size : 256;
  if (c->block % 64 != 0 || c->block > 256) {
    delete c;
    return XNH_E_INVALID;
  }
  c->lds_depth = cfg->lds_stack > 0 ? std::min(cfg->lds_stack, kMaxStack) : 16;
  c->use_graph = (cfg->flags & XNH_FLAG_NO_GRAPH) == 0;
  const long long npix = static_cast<long long>(cfg->width) * cfg->height;
  if (cfg->spp_per_launch > 0) {
    c->nsamples = cfg->spp_per_launch;
  } else {
    c->nsamples = default_nsamples(npix);
  }
  if (npix * c->nsamples > (1LL << 30)) {
    delete c;
    return XNH_E_INVALID;
  }
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return XNH_E_NO_DEVICE;
  }
  // The allocation-free entry point's static scratch: its device address,
  // once per device (xnh_image_noise_stats_noalloc).
  if (img_static_scratch_init(c->device) != hipSuccess) {
    delete c;
    return XNH_E_HIP;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev_begin) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess) {
    xnh_destroy(c);
    return XNH_E_HIP;
  }
  c->grid_trace = trace_grid(c->block, c->lds_depth);
  const int cus = num_cus();
  if (c->grid_trace <= 0 || cus <= 0) {
    xnh_destroy(c);
    return XNH_E_HIP;
  }
  c->grid_shade = cus * 8;
  int rc = alloc_pixels(c);
  if (rc == XNH_OK) rc = alloc_waves(c);
  if (rc != XNH_OK) {
    xnh_destroy(c);
    return rc;
  }
  *out = c;
  return XNH_OK;
}

void xnh_destroy(xnh_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->graph) hipGraphExecDestroy(c->graph);
  free_all(c->own.allocs);
  free_all(c->wave_allocs);
  free_all(c->pix_allocs);
  free_all(c->guide.allocs);
  free_all(c->out.allocs);
  if (c->d_jobs) hipFree(c->d_jobs);
  if (c->d_chunks) hipFree(c->d_chunks);
  if (c->d_nchunks) hipFree(c->d_nchunks);
  for (hipEvent_t e : c->prof_events) hipEventDestroy(e);
  if (c->ev_begin) hipEventDestroy(c->ev_begin);
  if (c->ev_end) hipEventDestroy(c->ev_end);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

const char* xnh_last_error(const xnh_ctx* c) { return c ? c->err.c_str() : "null context"; }

int xnh_upload_scene(xnh_ctx* c, const xnh_flat_scene* fs) {
  if (!c || !fs) return XNH_E_INVALID;
  c->has_scene = false;
  const int rc = flatten_upload(c, fs, &c->own);
  if (rc == XNH_OK) c->has_scene = true;
  return rc;
}

int xnh_scene_upload(xnh_ctx* c, const xnh_flat_scene* fs, xnh_gpu_scene** out) {
  if (!c || !fs || !out) return XNH_E_INVALID;
  *out = nullptr;
  xnh_gpu_scene* g = new (std::nothrow) xnh_gpu_scene();
  if (!g) return XNH_E_ALLOC;
  const int rc = flatten_upload(c, fs, g);
  if (rc != XNH_OK) {
    delete g;
    return rc;
  }
  *out = g;
  return XNH_OK;
}

void xnh_scene_release(xnh_ctx* c, xnh_gpu_scene* g) {
  if (!g) return;
  if (c && c->stream) hipStreamSynchronize(c->stream);
  free_all(g->allocs);
  delete g;
}

int xnh_set_profiling(xnh_ctx* c, int32_t flags) {
  if (!c || (flags & ~(XNH_PROFILE_TIMING | XNH_PROFILE_WORK)) != 0) return XNH_E_INVALID;
  c->profiling = (flags & XNH_PROFILE_TIMING) != 0;
  const bool work = (flags & XNH_PROFILE_WORK) != 0;
  if (work != c->counting) {
    c->counting = work;
    c->table_gen++;   // the captured graph holds the other kernel variants
  }
  return XNH_OK;
}

int xnh_render(xnh_ctx* c, const xnh_camera* cam, int32_t spp_first, int32_t spp_count, float* rgb_out) {
  if (!c || !cam || spp_count < 0 || spp_first < 0) return XNH_E_INVALID;
  if (!c->has_scene) return fail(c, XNH_E_NO_SCENE, "xnh_render before xnh_upload_scene");
  const int npix = c->cfg.width * c->cfg.height;
  std::vector<JobSpec> jobs;
  if (spp_count > 0)
    jobs.push_back(JobSpec{&c->own.S, *cam, c->cfg.width, c->cfg.height, spp_first, spp_count, c->cfg.max_bounces,
                           c->cfg.min_bounces, c->accum});
  const int rc = render_jobs(c, jobs);
  if (rc != XNH_OK) return rc;
  c->spp_total += spp_count;
  if (rgb_out && c->spp_total > 0) {
    const float inv = 1.0f / static_cast<float>(c->spp_total);
    HIP_OR_FAIL(c, launch_average(c->stream, c->accum, c->d_out, npix, inv));
    HIP_OR_FAIL(c, hipMemcpyAsync(rgb_out, c->d_out, sizeof(float) * 3 * npix, hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  }
  return XNH_OK;
}

int xnh_render_jobs(xnh_ctx* c, const xnh_job* jobs, int32_t njobs) {
  if (!c || njobs < 0 || (njobs > 0 && !jobs)) return XNH_E_INVALID;
  std::vector<JobSpec> spec;
  std::vector<void*> tmp;
  std::vector<float4*> acc(njobs, nullptr);
  for (int j = 0; j < njobs; ++j) {
    const xnh_job& jb = jobs[j];
    if (jb.width <= 0 || jb.height <= 0 || jb.spp_first < 0 || jb.spp_count < 0 || jb.max_bounces < 0 ||
        jb.max_bounces > 64 || jb.min_bounces < 0) {
      free_all(tmp);
      return fail(c, XNH_E_INVALID, "bad job");
    }
    const DevScene* S = jb.scene ? &jb.scene->S : (c->has_scene ? &c->own.S : nullptr);
    if (!S) {
      free_all(tmp);
      return fail(c, XNH_E_NO_SCENE, "job without a scene");
    }
    const size_t npix = static_cast<size_t>(jb.width) * jb.height;
    int rc = dev_alloc(c, tmp, &acc[j], npix);
    if (rc != XNH_OK) {
      free_all(tmp);
      return rc;
    }
    if (hipMemsetAsync(acc[j], 0, sizeof(float4) * npix, c->stream) != hipSuccess) {
      free_all(tmp);
      return fail(c, XNH_E_HIP, "hipMemsetAsync");
    }
    spec.push_back(JobSpec{S, jb.camera, jb.width, jb.height, jb.spp_first, jb.spp_count, jb.max_bounces,
                           jb.min_bounces, acc[j]});
  }
  int rc = render_jobs(c, spec);
  for (int j = 0; j < njobs && rc == XNH_OK; ++j) {
    const xnh_job& jb = jobs[j];
    const size_t npix = static_cast<size_t>(jb.width) * jb.height;
    if (!jb.rgb_out && !jb.sum_out) continue;
    std::vector<float4> h(npix);
    if (hipMemcpy(h.data(), acc[j], sizeof(float4) * npix, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(c, XNH_E_HIP, "hipMemcpy");
      break;
    }
    const float inv = jb.spp_count > 0 ? 1.0f / static_cast<float>(jb.spp_count) : 0.f;
    for (size_t p = 0; p < npix; ++p) {
      if (jb.sum_out) {
        jb.sum_out[3 * p] = h[p].x;
        jb.sum_out[3 * p + 1] = h[p].y;
        jb.sum_out[3 * p + 2] = h[p].z;
      }
      if (jb.rgb_out) {
        jb.rgb_out[3 * p] = h[p].x * inv;
        jb.rgb_out[3 * p + 1] = h[p].y * inv;
        jb.rgb_out[3 * p + 2] = h[p].z * inv;
      }
    }
  }
  hipStreamSynchronize(c->stream);
  free_all(tmp);
  return rc;
}

namespace {

int ensure_out(xnh_ctx* c, int64_t npix) {
  if (npix <= c->out.cap) return XNH_OK;
  free_all(c->out.allocs);
  c->out.cap = 0;
  auto& o = c->out;
  int rc = dev_alloc(c, o.allocs, &o.img, static_cast<size_t>(3 * npix));
  if (rc == XNH_OK) rc = dev_alloc(c, o.allocs, &o.rgb8, static_cast<size_t>(3 * npix));
  if (rc == XNH_OK) rc = dev_alloc(c, o.allocs, &o.sc.lum, static_cast<size_t>(npix));
  if (rc == XNH_OK) rc = dev_alloc(c, o.allocs, &o.sc.key, static_cast<size_t>(npix));
  if (rc == XNH_OK) rc = dev_alloc(c, o.allocs, &o.sc.part, static_cast<size_t>(2 * kImgMaxBlocks));
  if (rc == XNH_OK) rc = dev_alloc(c, o.allocs, &o.sc.pmax, static_cast<size_t>(kImgMaxBlocks));
  if (rc == XNH_OK) rc = dev_alloc(c, o.allocs, &o.sc.pcnt, static_cast<size_t>(kImgMaxBlocks));
  if (rc == XNH_OK) rc = dev_alloc(c, o.allocs, &o.sc.st, 1);
  if (rc != XNH_OK) {
    free_all(o.allocs);
    return rc;
  }
  o.cap = npix;
  return XNH_OK;
}

// The framebuffer of an output-stage call in c->out.img: the caller's host
// image, or the running average in the swapchain layout.
int output_source(xnh_ctx* c, const float* rgb, int w, int h) {
  const int64_t npix = static_cast<int64_t>(w) * h;
  if (npix > kImgMaxPixels) return fail(c, XNH_E_INVALID, "image larger than 2^26 pixels");
  if (!rgb && (w != c->cfg.width || h != c->cfg.height))
    return fail(c, XNH_E_INVALID, "the running average has the context's size");
  const int rc = ensure_out(c, npix);
  if (rc != XNH_OK) return rc;
  if (rgb) {
    HIP_OR_FAIL(c, hipMemcpyAsync(c->out.img, rgb, sizeof(float) * 3 * npix, hipMemcpyHostToDevice, c->stream));
  } else {
    const float inv = c->spp_total > 0 ? 1.0f / static_cast<float>(c->spp_total) : 0.f;
    HIP_OR_FAIL(c, launch_img_from_accum(c->stream, c->accum, inv, w, h, c->out.img));
  }
  return XNH_OK;
}

}  // namespace

namespace {

// Reads the device state of a finished NoiseStats computation into `out`
// (math/image.h:100-118).
int read_noise_stats(xnh_ctx* c, const ImgState* dst, int32_t w, int32_t h, xnh_noise_stats* out) {
  ImgState st;
  HIP_OR_FAIL(c, hipMemcpyAsync(&st, dst, sizeof st, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  const int interior = (w - 2) * (h - 2);
  out->total_interior_pixel_count = interior;
  out->avg_luminance = st.avg;
  out->edge_threshold = st.edge;
  out->epsilon = st.eps;
  out->valid_pixel_count = static_cast<int32_t>(st.count);
  out->valid_pixel_ratio = static_cast<float>(st.count) / static_cast<float>(interior);
  out->max_absolute = st.max_abs;
  if (st.count > 0) {   // image.h:108-115
    out->mean_squared_score = st.mean_sq;
    out->rms_score = std::sqrt(st.mean_sq);
    out->mean_absolute = st.mean_abs;
    std::memcpy(&out->median_absolute, &st.prefix[0], 4);
    std::memcpy(&out->p95_absolute, &st.prefix[1], 4);
  }
  return XNH_OK;
}

}  // namespace

int xnh_image_noise_stats(xnh_ctx* c, const float* rgb, int32_t w, int32_t h, xnh_noise_stats* out) {
  if (!c || !out || w < 0 || h < 0) return XNH_E_INVALID;
  *out = xnh_noise_stats{};
  if (w <= 2 || h <= 2) return XNH_OK;   // image.h:41-42
  int rc = output_source(c, rgb, w, h);
  if (rc != XNH_OK) return rc;
  HIP_OR_FAIL(c, launch_noise_stats(c->stream, c->out.img, w, h, c->out.sc));
  return read_noise_stats(c, c->out.sc.st, w, h, out);
}

int xnh_image_noise_stats_noalloc(xnh_ctx* c, const float* rgb_dev, int32_t w, int32_t h, xnh_noise_stats* out) {
  if (!c || !out || w < 0 || h < 0) return XNH_E_INVALID;
  *out = xnh_noise_stats{};
  if (w <= 2 || h <= 2) return XNH_OK;   // image.h:41-42
  if (!rgb_dev) return fail(c, XNH_E_INVALID, "no image");
  const int64_t npix = static_cast<int64_t>(w) * h;
  if (npix > kImgStaticMaxPixels)
    return fail(c, XNH_E_INVALID, "image larger than the static scratch (2^22 pixels)");
  // The image must lie inside one device allocation: the kernels read
  // 3 * w * h floats from rgb_dev unchecked.
  hipDeviceptr_t base = nullptr;
  size_t bytes = 0;
  const hipError_t range = hipMemGetAddressRange(&base, &bytes, const_cast<float*>(rgb_dev));
  if (range != hipSuccess) {
    (void)hipGetLastError();   // not device memory: clear the error the lookup left
    return fail(c, XNH_E_INVALID, "the image is not in device memory");
  }
  const uintptr_t b = reinterpret_cast<uintptr_t>(base), p = reinterpret_cast<uintptr_t>(rgb_dev);
  if (p < b || p - b + sizeof(float) * 3 * static_cast<size_t>(npix) > bytes)
    return fail(c, XNH_E_INVALID, "the image extends past the end of its device allocation");
  ImgScratch sc{};
  hipError_t e = img_static_scratch(c->device, &sc);
  if (e == hipErrorNotInitialized) {   // set-up did not run on this device (xnh_create does it)
    e = img_static_scratch_init(c->device);
    if (e == hipSuccess) e = img_static_scratch(c->device, &sc);
  }
  if (e != hipSuccess) return fail(c, XNH_E_HIP, std::string("static scratch: ") + hipGetErrorString(e));
  HIP_OR_FAIL(c, launch_noise_stats(c->stream, rgb_dev, w, h, sc));
  return read_noise_stats(c, sc.st, w, h, out);
}

int xnh_image_rgb8(xnh_ctx* c, const float* rgb, int32_t w, int32_t h, uint8_t* out) {
  if (!c || !out || w < 0 || h < 0) return XNH_E_INVALID;
  if (w == 0 || h == 0) return XNH_OK;
  int rc = output_source(c, rgb, w, h);
  if (rc != XNH_OK) return rc;
  HIP_OR_FAIL(c, launch_img_rgb8(c->stream, c->out.img, w, h, c->out.rgb8));
  HIP_OR_FAIL(c, hipMemcpyAsync(out, c->out.rgb8, static_cast<size_t>(3) * w * h, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return XNH_OK;
}

namespace {

// Vertex indices in range (the device kernels index with them unchecked).
int check_mesh_indices(xnh_ctx* c, const xnh_mesh* m) {
  const int n = m->num_triangles, nv = m->num_vertices;
  if (n < 0 || nv < 0 || (n > 0 && (!m->vx || !m->vy || !m->vz || !m->indices)))
    return fail(c, XNH_E_INVALID, "mesh arrays missing");
  if (n > (1 << 28)) return fail(c, XNH_E_INVALID, "more than 2^28 triangles");
  for (int64_t k = 0; k < 3 * static_cast<int64_t>(n); ++k)
    if (m->indices[k] < 0 || m->indices[k] >= nv) return fail(c, XNH_E_INVALID, "vertex index out of range");
  return XNH_OK;
}

struct DevMesh {   // one mesh's vertices and indices on the device, freed on scope exit
  std::vector<void*> allocs;
  float *vx = nullptr, *vy = nullptr, *vz = nullptr;
  int32_t* idx = nullptr;
  ~DevMesh() { free_all(allocs); }
};

int upload_mesh(xnh_ctx* c, const xnh_mesh* m, DevMesh& d) {
  const size_t nv = static_cast<size_t>(m->num_vertices), ni = 3 * static_cast<size_t>(m->num_triangles);
  int rc = dev_alloc(c, d.allocs, &d.vx, nv);
  if (rc == XNH_OK) rc = dev_alloc(c, d.allocs, &d.vy, nv);
  if (rc == XNH_OK) rc = dev_alloc(c, d.allocs, &d.vz, nv);
  if (rc == XNH_OK) rc = dev_alloc(c, d.allocs, &d.idx, ni);
  if (rc != XNH_OK) return rc;
  HIP_OR_FAIL(c, hipMemcpyAsync(d.vx, m->vx, sizeof(float) * nv, hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(c, hipMemcpyAsync(d.vy, m->vy, sizeof(float) * nv, hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(c, hipMemcpyAsync(d.vz, m->vz, sizeof(float) * nv, hipMemcpyHostToDevice, c->stream));
  HIP_OR_FAIL(c, hipMemcpyAsync(d.idx, m->indices, sizeof(int32_t) * ni, hipMemcpyHostToDevice, c->stream));
  return XNH_OK;
}

}  // namespace

int xnh_build_blas(xnh_ctx* c, const xnh_mesh* m, xnh_bvh_node* nodes_out, uint32_t* prims_out,
                   int32_t* num_nodes_out) {
  if (!c || !m || !num_nodes_out) return XNH_E_INVALID;
  int rc = check_mesh_indices(c, m);
  if (rc != XNH_OK) return rc;
  const int n = m->num_triangles;
  *num_nodes_out = n > 0 ? 2 * n - 1 : 0;
  if (n == 0) return XNH_OK;
  if (!nodes_out || !prims_out) return XNH_E_INVALID;
  if (n == 1) {   // a single leaf: its box is the triangle's
    xnh_bvh_node v{};
    for (int k = 0; k < 3; ++k) {
      const float* a = k == 0 ? m->vx : (k == 1 ? m->vy : m->vz);
      const float p0 = a[m->indices[0]], p1 = a[m->indices[1]], p2 = a[m->indices[2]];
      v.lo[k] = std::fmin(std::fmin(p0, p1), p2);
      v.hi[k] = std::fmax(std::fmax(p0, p1), p2);
    }
    v.child = 0;
    v.count = 1;
    nodes_out[0] = v;
    prims_out[0] = 0;
    return XNH_OK;
  }
  DevMesh d;
  rc = upload_mesh(c, m, d);
  if (rc != XNH_OK) return rc;
  std::vector<void*> own;
  xnh_bvh_node* nodes = nullptr;
  uint32_t *prims = nullptr, *err = nullptr;
  rc = dev_alloc(c, own, &nodes, static_cast<size_t>(2 * n - 1));
  if (rc == XNH_OK) rc = dev_alloc(c, own, &prims, static_cast<size_t>(n));
  if (rc == XNH_OK) rc = dev_alloc(c, own, &err, 1);
  uint32_t herr = 0;
  hipError_t e = hipSuccess;
  if (rc == XNH_OK) {
    e = hipMemsetAsync(err, 0, 4, c->stream);
    if (e == hipSuccess) e = build_blas_device(c->stream, d.vx, d.vy, d.vz, d.idx, n, nodes, prims, err);
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(nodes_out, nodes, sizeof(xnh_bvh_node) * (2 * n - 1), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(prims_out, prims, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream);
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = es;
  }
  free_all(own);
  if (rc != XNH_OK) return rc;
  if (e != hipSuccess) return fail(c, XNH_E_HIP, std::string("device BLAS build: ") + hipGetErrorString(e));
  if (herr != 0) return fail(c, XNH_E_HIP, "device BLAS build produced a malformed tree (mask " + std::to_string(herr) + ")");
  return XNH_OK;
}

int xnh_refit_blas(xnh_ctx* c, const xnh_mesh* m, xnh_bvh_node* nodes_out) {
  if (!c || !m || !nodes_out) return XNH_E_INVALID;
  int rc = check_mesh_indices(c, m);
  if (rc != XNH_OK) return rc;
  const int nn = m->num_nodes, n = m->num_triangles;
  if (nn <= 0) return XNH_OK;
  if (!m->nodes || !m->prims) return fail(c, XNH_E_INVALID, "mesh has no BLAS");
  // the kernels trust the structure: children after their parent, leaves
  // inside the slot array, slots naming triangles
  for (int j = 0; j < nn; ++j) {
    const xnh_bvh_node& v = m->nodes[j];
    if (v.count == 0 && (v.child <= static_cast<uint32_t>(j) || static_cast<int64_t>(v.child) + 1 >= nn))
      return fail(c, XNH_E_INVALID, "BLAS child outside node range");
    if (v.count > 0 && static_cast<int64_t>(v.child) + v.count > n)
      return fail(c, XNH_E_INVALID, "BLAS leaf outside slot range");
  }
  for (int k = 0; k < n; ++k)
    if (m->prims[k] >= static_cast<uint32_t>(n)) return fail(c, XNH_E_INVALID, "slot names no triangle");
  DevMesh d;
  rc = upload_mesh(c, m, d);
  if (rc != XNH_OK) return rc;
  std::vector<void*> own;
  xnh_bvh_node* nodes = nullptr;
  uint32_t *prims = nullptr, *parent = nullptr, *flag = nullptr, *err = nullptr;
  rc = dev_alloc(c, own, &nodes, static_cast<size_t>(nn));
  if (rc == XNH_OK) rc = dev_alloc(c, own, &prims, static_cast<size_t>(n));
  if (rc == XNH_OK) rc = dev_alloc(c, own, &parent, static_cast<size_t>(nn));
  if (rc == XNH_OK) rc = dev_alloc(c, own, &flag, static_cast<size_t>(nn));
  if (rc == XNH_OK) rc = dev_alloc(c, own, &err, 1);
  uint32_t herr = 0;
  hipError_t e = hipSuccess;
  if (rc == XNH_OK) {
    e = hipMemcpyAsync(nodes, m->nodes, sizeof(xnh_bvh_node) * nn, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(prims, m->prims, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(err, 0, 4, c->stream);
    if (e == hipSuccess) e = refit_blas_device(c->stream, nodes, nn, prims, n, d.vx, d.vy, d.vz, d.idx, parent, flag, err);
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(nodes_out, nodes, sizeof(xnh_bvh_node) * nn, hipMemcpyDeviceToHost, c->stream);
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = es;
  }
  free_all(own);
  if (rc != XNH_OK) return rc;
  if (e != hipSuccess) return fail(c, XNH_E_HIP, std::string("device BLAS refit: ") + hipGetErrorString(e));
  if (herr != 0) return fail(c, XNH_E_HIP, "device BLAS refit found a malformed tree (mask " + std::to_string(herr) + ")");
  return XNH_OK;
}

int xnh_read_accum(xnh_ctx* c, float* sum_out, int32_t* spp_out) {
  if (!c) return XNH_E_INVALID;
  const int npix = c->cfg.width * c->cfg.height;
  if (sum_out) {
    std::vector<float4> tmp(npix);
    HIP_OR_FAIL(c, hipMemcpyAsync(tmp.data(), c->accum, sizeof(float4) * npix, hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    for (int p = 0; p < npix; ++p) {
      sum_out[3 * p] = tmp[p].x;
      sum_out[3 * p + 1] = tmp[p].y;
      sum_out[3 * p + 2] = tmp[p].z;
    }
  }
  if (spp_out) *spp_out = c->spp_total;
  return XNH_OK;
}

int xnh_reset(xnh_ctx* c) {
  if (!c) return XNH_E_INVALID;
  HIP_OR_FAIL(c, hipMemsetAsync(c->accum, 0, sizeof(float4) * c->cfg.width * c->cfg.height, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  c->spp_total = 0;
  return XNH_OK;
}

// ---- tuning -------------------------------------------------------------------

int xnh_set_tuning(xnh_ctx* c, int32_t block_size, int32_t lds_stack, int32_t spp_per_launch) {
  if (!c) return XNH_E_INVALID;
  const int block = block_size > 0 ? block_size : 256;
  const int lds = lds_stack > 0 ? lds_stack : 16;
  if (block % 64 != 0 || block > 256 || lds > kMaxStack || spp_per_launch < 0)
    return fail(c, XNH_E_INVALID, "tuning out of range (block 64..256 in steps of 64, lds_stack 1..128)");
  const long long npix = static_cast<long long>(c->cfg.width) * c->cfg.height;
  int ns = spp_per_launch;
  if (ns == 0) ns = default_nsamples(npix);
  if (npix * ns > (1LL << 30)) return fail(c, XNH_E_INVALID, "queue capacity above 2^30 paths");
  if (c->guide.on && ns != c->nsamples) return fail(c, XNH_E_INVALID, "queue capacity is fixed while guiding is on");
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  const int grid = trace_grid(block, lds);
  if (grid <= 0) return fail(c, XNH_E_HIP, "occupancy query failed");
  // from here on the queues are reallocated and the accumulation cleared,
  // whether or not the new tuning fits
  const int old_block = c->block, old_lds = c->lds_depth, old_ns = c->nsamples, old_grid = c->grid_trace;
  c->table_gen++;   // the captured graph holds the old buffers
  c->spp_total = 0;
  HIP_OR_FAIL(c, hipMemset(c->accum, 0, sizeof(float4) * static_cast<size_t>(npix)));
  auto apply = [&](int b, int l, int n, int g) {
    c->block = b;
    c->lds_depth = l;
    c->nsamples = n;
    c->grid_trace = g;
    const int r = alloc_waves(c);
    attach_guide(c);
    return r;
  };
  const int rc = apply(block, lds, ns, grid);
  if (rc == XNH_OK) return XNH_OK;
  const std::string why = c->err;
  if (apply(old_block, old_lds, old_ns, old_grid) != XNH_OK)
    return fail(c, rc, why + "; restoring the previous tuning failed too (" + c->err +
                           "): the context renders nothing until a later xnh_set_tuning succeeds");
  return fail(c, rc, why + " (the previous tuning is kept)");
}

int xnh_get_tuning(const xnh_ctx* c, int32_t* block_size, int32_t* lds_stack, int32_t* spp_per_launch) {
  if (!c) return XNH_E_INVALID;
  if (block_size) *block_size = c->block;
  if (lds_stack) *lds_stack = c->lds_depth;
  if (spp_per_launch) *spp_per_launch = c->nsamples;
  return XNH_OK;
}

int xnh_autotune(xnh_ctx* c, const xnh_camera* cam, int32_t probe_spp, xnh_tune_result* out) {
  if (!c || !cam || probe_spp <= 0) return XNH_E_INVALID;
  if (!c->has_scene) return fail(c, XNH_E_NO_SCENE, "xnh_autotune before xnh_upload_scene");
  if (c->B.capacity == 0) return fail(c, XNH_E_ALLOC, "no wavefront queues (a re-tuning failed)");
  const long long npix = static_cast<long long>(c->cfg.width) * c->cfg.height;
  const int block0 = c->block, lds0 = c->lds_depth, ns0 = c->nsamples;
  // samples per batch: half, the configured value (the default, about 2^28
  // paths, unless set) and double; fixed while guiding (the guide's
  // per-entry buffers are sized by the queues)
  std::vector<int> caps = {ns0};
  if (!c->guide.on) {
    if (ns0 / 2 >= 1) caps.push_back(ns0 / 2);
    if (npix * ns0 * 2 <= (1LL << 30) && ns0 * 2 <= 256) caps.push_back(ns0 * 2);
  }
  const int blocks[] = {64, 128, 256};
  const int ldss[] = {8, 12, 16, 24, 32};
  float best = 3.0e38f;
  int bb = block0, bl = lds0, bn = ns0, tried = 0;
  const bool prof = c->profiling;
  const bool train = c->guide.train;
  c->profiling = false;
  c->guide.train = false;   // probes leave the guide's parameters as they are
  // Each candidate renders one untimed batch, then `probe` samples, at least
  // two full batches of the stage's largest candidate, so that a larger
  // batch is not timed on a partial one.
  auto time_one = [&](int b, int l, int ns, int probe) {
    if (xnh_set_tuning(c, b, l, ns) != XNH_OK) return;
    int rc = xnh_render(c, cam, 1, ns, nullptr);
    if (rc == XNH_OK) rc = xnh_render(c, cam, 1, probe, nullptr);
    if (rc != XNH_OK) return;
    ++tried;
    const float ms = static_cast<float>(c->stats.render_ms) / static_cast<float>(probe);
    if (ms < best) {
      best = ms;
      bb = b;
      bl = l;
      bn = ns;
    }
  };
  // Stage 1: samples per batch at the configured workgroup size and LDS
  // depth. Stage 2: workgroup size and LDS depth at the chosen batch. (The
  // batch size moves throughput by up to 2x, the other two by about 1 %:
  // searching them one after the other costs 3 + 15 probes instead of 45.)
  const int ns_max = *std::max_element(caps.begin(), caps.end());
  for (int ns : caps) time_one(block0, lds0, ns, std::max(probe_spp, 2 * ns_max));
  const int ns_best = bn;
  best = 3.0e38f;   // stage 2 times another probe length: compare within it
  for (int b : blocks)
    for (int l : ldss) time_one(b, l, ns_best, std::max(probe_spp, 2 * ns_best));
  c->profiling = prof;
  c->guide.train = train;
  const int rc = xnh_set_tuning(c, bb, bl, bn);
  if (rc != XNH_OK) return rc;
  if (tried == 0) return fail(c, XNH_E_HIP, "no tuning candidate rendered");
  if (out) {
    out->block_size = bb;
    out->lds_stack = bl;
    out->spp_per_launch = bn;
    out->ms_per_spp = best;
    out->candidates = tried;
  }
  return XNH_OK;
}

int xnh_spp(const xnh_ctx* c) { return c ? c->spp_total : XNH_E_INVALID; }

int xnh_get_stats(const xnh_ctx* c, xnh_stats* out) {
  if (!c || !out) return XNH_E_INVALID;
  *out = c->stats;
  return XNH_OK;
}

// ---- path guiding -------------------------------------------------------------

static uint64_t mix64(uint64_t& s) {   // xorshift64* for parameter initialisation
  s ^= s >> 12;
  s ^= s << 25;
  s ^= s >> 27;
  return s * 2685821657736338717ULL;
}

int xnh_guide_disable(xnh_ctx* c) {
  if (!c) return XNH_E_INVALID;
  if (c->stream) hipStreamSynchronize(c->stream);
  free_all(c->guide.allocs);
  c->guide = GuideState{};
  attach_guide(c);
  c->table_gen++;
  return XNH_OK;
}

int xnh_guide_set_training(xnh_ctx* c, int32_t on) {
  if (!c || !c->guide.on) return XNH_E_INVALID;
  c->guide.train = on != 0;
  return XNH_OK;
}

int xnh_guide_enable(xnh_ctx* c, const xnh_guide_config* cfg_in) {
  if (!c || !cfg_in) return XNH_E_INVALID;
  xnh_guide_config cfg = *cfg_in;
  if (cfg.t_log2 == 0) cfg.t_log2 = 19;
  if (cfg.learning_rate <= 0.f) cfg.learning_rate = 5e-3f;
  if (cfg.guide_fraction <= 0.f) cfg.guide_fraction = 0.5f;
  if (cfg.train_batch <= 0) cfg.train_batch = 1 << 18;
  if (c->B.capacity == 0) return fail(c, XNH_E_ALLOC, "no wavefront queues (a re-tuning failed)");
  const size_t gcap = guided_cap(c);
  if (cfg.record_paths <= 0 || static_cast<size_t>(cfg.record_paths) > gcap) cfg.record_paths = static_cast<int32_t>(gcap);
  // one optimisation step per 2^22 paths a guided batch holds: the guide
  // then learns at the same rate per rendered sample whatever the batch
  // size. Caustic room, 2^26-path batches, equal time against plain: 1
  // step per batch 1.06x (trimmed relMSE) / 1.08x (median pixel error),
  // 8 steps 1.12x / 1.33x, 16 steps 1.13x / 1.36x, 32 steps 1.14x / 1.37x.
  if (cfg.train_steps <= 0)
    cfg.train_steps = std::max(
        1, std::min(64, static_cast<int>(std::min<size_t>(static_cast<size_t>(c->B.capacity), kGuidedBatchPaths) >> 22)));
  if (cfg.warmup_steps < 0) cfg.warmup_steps = 0;
  if (cfg.guide_depth <= 0) cfg.guide_depth = 1;
  if (cfg.t_log2 < 8 || cfg.t_log2 > 24 || cfg.guide_fraction >= 1.f) return fail(c, XNH_E_INVALID, "guide config");
  xnh_guide_disable(c);
  GuideState& g = c->guide;
  g.cfg = cfg;
  const size_t T = size_t(1) << cfg.t_log2;
  g.ngrid = size_t(kGLevels) * T * kGFeat;
  int rc = dev_alloc(c, g.allocs, &g.grid_m, g.ngrid);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.grid_g, g.ngrid);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.grid_am, g.ngrid);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.grid_av, g.ngrid);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.grid_h, g.ngrid);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.w_m, size_t(kWTotal));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.w_g, size_t(kWTotal));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.w_am, size_t(kWTotal));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.w_av, size_t(kWTotal));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.w_h, size_t(kWTotal));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.w_t, size_t(kWTotal));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.loss, size_t(1));
  const size_t R = static_cast<size_t>(cfg.train_batch);
  // records: one per guided vertex of a recording path, at most 2^25 a
  // batch. The recording paths are a hashed 1-in-rec_every choice of the
  // batch's paths, rec_every the smallest that keeps their expected number
  // within that (so the records cover the whole image rather than its
  // first-shaded part). kQSeg equal segments, each with room over an equal
  // share for one shading window's entries per guided depth and 1/64 for
  // the draw's spread, so no record is dropped in practice.
  const size_t rp = std::min(static_cast<size_t>(cfg.record_paths),
                             (static_cast<size_t>(1) << 25) / static_cast<size_t>(cfg.guide_depth));
  size_t RC = rp * static_cast<size_t>(cfg.guide_depth);
  RC += RC / 64 + static_cast<size_t>(kQSeg) * kShadeWin * 64 * static_cast<size_t>(cfg.guide_depth);
  RC = (RC + kQSeg - 1) / kQSeg * kQSeg;
  // per queue entry of a guided batch: its slots and the segment slack
  const size_t gq = std::min(static_cast<size_t>(c->B.qcap), gcap + static_cast<size_t>(kQueueSlack));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.mix, size_t(kMixVecs) * gq);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.feat, size_t(6) * gq);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.rec, 4 * RC);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.lam_all, RC);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.lam_sum_all, size_t(1));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.t_pos, 3 * R);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.t_nrm, 3 * R);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.t_dir, 3 * R);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.t_lam, R);
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.t_lam_sum, size_t(1));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.t_n, size_t(1));
  if (rc == XNH_OK) rc = dev_alloc(c, g.allocs, &g.alpha_dev, size_t(1));
  if (rc != XNH_OK) {
    xnh_guide_disable(c);
    return rc;
  }
  g.net.grid = g.grid_h;
  g.net.w = g.w_h;
  g.net.wt = g.w_t;
  g.mix_stride = static_cast<uint32_t>(gq);
  g.slot_cap = static_cast<uint32_t>(gcap);
  g.rec_cap = static_cast<uint32_t>(RC);
  g.rec_every = static_cast<uint32_t>(std::max<size_t>(1, (gcap + rp - 1) / rp));
  g.alpha_host = cfg.warmup_steps > 0 ? 0.f : cfg.guide_fraction;
  if (hipMemcpy(g.alpha_dev, &g.alpha_host, sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    xnh_guide_disable(c);
    return fail(c, XNH_E_HIP, "guide alpha upload");
  }
  c->table_gen++;   // the captured batch graph gains the guide kernels
  g.net.t_log2 = cfg.t_log2;
  const double b = std::exp((std::log(2048.0) - std::log(16.0)) / (kGLevels - 1));
  for (int l = 0; l < kGLevels; ++l) g.net.res[l] = static_cast<int32_t>(std::floor(16.0 * std::pow(b, l)));
  for (int k = 0; k < 3; ++k) {
    g.net.lo[k] = cfg.bounds_lo[k];
    const float e = cfg.bounds_hi[k] - cfg.bounds_lo[k];
    g.net.inv_ext[k] = e > 0.f ? 1.f / e : 1.f;
  }
  // initial parameters: grid U(-1e-4, 1e-4), weights Glorot-uniform
  std::vector<float> grid(g.ngrid), w(kWTotal);
  uint64_t st = 0x9E3779B97F4A7C15ULL ^ (uint64_t(cfg.seed) + 1);
  auto uni = [&]() { return static_cast<float>((mix64(st) >> 40) * (1.0 / 16777216.0)) * 2.f - 1.f; };
  for (auto& x : grid) x = 1e-4f * uni();
  const int dims[4][2] = {{kGIn, kGHid}, {kGHid, kGHid}, {kGHid, kGHid}, {kGHid, kGOut}};
  const int offs[4] = {kW1, kW2, kW3, kW4};
  for (int L = 0; L < 4; ++L) {
    const float a = std::sqrt(6.f / float(dims[L][0] + dims[L][1]));
    for (int i = 0; i < dims[L][0] * dims[L][1]; ++i) w[offs[L] + i] = a * uni();
  }
  g.on = true;
  attach_guide(c);
  return xnh_guide_set_params(c, grid.data(), w.data());
}

int xnh_guide_set_params(xnh_ctx* c, const float* grid, const float* w) {
  if (!c || !c->guide.on || !grid || !w) return XNH_E_INVALID;
  GuideState& g = c->guide;
  hipStream_t s = c->stream;
  HIP_OR_FAIL(c, hipMemcpyAsync(g.grid_m, grid, g.ngrid * 4, hipMemcpyHostToDevice, s));
  HIP_OR_FAIL(c, hipMemcpyAsync(g.w_m, w, size_t(kWTotal) * 4, hipMemcpyHostToDevice, s));
  HIP_OR_FAIL(c, launch_to_half(s, g.grid_m, g.grid_h, g.ngrid));
  HIP_OR_FAIL(c, launch_to_half(s, g.w_m, g.w_h, size_t(kWTotal)));
  HIP_OR_FAIL(c, launch_tile_weights(s, g.w_h, g.w_t));
  for (float* p : {g.grid_g, g.grid_am, g.grid_av}) HIP_OR_FAIL(c, hipMemsetAsync(p, 0, g.ngrid * 4, s));
  for (float* p : {g.w_g, g.w_am, g.w_av}) HIP_OR_FAIL(c, hipMemsetAsync(p, 0, size_t(kWTotal) * 4, s));
  HIP_OR_FAIL(c, hipStreamSynchronize(s));
  g.step = 0;
  return XNH_OK;
}

int xnh_guide_get_params(xnh_ctx* c, float* grid, float* w) {
  if (!c || !c->guide.on) return XNH_E_INVALID;
  GuideState& g = c->guide;
  if (grid) HIP_OR_FAIL(c, hipMemcpyAsync(grid, g.grid_m, g.ngrid * 4, hipMemcpyDeviceToHost, c->stream));
  if (w) HIP_OR_FAIL(c, hipMemcpyAsync(w, g.w_m, size_t(kWTotal) * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
  return XNH_OK;
}

int xnh_guide_eval(xnh_ctx* c, int32_t n, const float* pos, const float* nrm, float* raw) {
  if (!c || !c->guide.on || n < 0 || (n > 0 && (!pos || !nrm || !raw))) return XNH_E_INVALID;
  if (n == 0) return XNH_OK;
  std::vector<void*> tmp;
  float *dp, *dn, *dr;
  int rc = dev_alloc(c, tmp, &dp, size_t(n) * 3);
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &dn, size_t(n) * 3);
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &dr, size_t(n) * kGOut);
  hipError_t e = hipSuccess;
  if (rc == XNH_OK) {
    e = hipMemcpyAsync(dp, pos, size_t(n) * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dn, nrm, size_t(n) * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_guide_eval(c->stream, c->guide.net, n, dp, dn, dr);
    if (e == hipSuccess) e = hipMemcpyAsync(raw, dr, size_t(n) * kGOut * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  free_all(tmp);
  if (rc != XNH_OK) return rc;
  if (e != hipSuccess) return fail(c, XNH_E_HIP, std::string("guide eval: ") + hipGetErrorString(e));
  return XNH_OK;
}

int xnh_guide_query(xnh_ctx* c, int32_t n, const float* pos, const float* nrm, const float* u, float* dir,
                    float* pdf) {
  if (!c || !c->guide.on || n < 0 || (n > 0 && (!pos || !nrm || !dir || !pdf))) return XNH_E_INVALID;
  if (n == 0) return XNH_OK;
  std::vector<void*> tmp;
  float *dp, *dn, *du = nullptr, *dd, *dq;
  int rc = dev_alloc(c, tmp, &dp, size_t(n) * 3);
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &dn, size_t(n) * 3);
  if (rc == XNH_OK && u) rc = dev_alloc(c, tmp, &du, size_t(n) * 3);
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &dd, size_t(n) * 3);
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &dq, size_t(n));
  hipError_t e = hipSuccess;
  if (rc == XNH_OK) {
    hipStream_t s = c->stream;
    e = hipMemcpyAsync(dp, pos, size_t(n) * 12, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dn, nrm, size_t(n) * 12, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && u) e = hipMemcpyAsync(du, u, size_t(n) * 12, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !u) e = hipMemcpyAsync(dd, dir, size_t(n) * 12, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_guide_query(s, c->guide.net, n, dp, dn, du, dd, dq);
    if (e == hipSuccess && u) e = hipMemcpyAsync(dir, dd, size_t(n) * 12, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(pdf, dq, size_t(n) * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  free_all(tmp);
  if (rc != XNH_OK) return rc;
  if (e != hipSuccess) return fail(c, XNH_E_HIP, std::string("guide query: ") + hipGetErrorString(e));
  return XNH_OK;
}

// One step on device-resident samples (lam already on the device).
static int guide_step(xnh_ctx* c, int n, const float* dp, const float* dn, const float* dd, const float* dl,
                      float lam_scale, bool apply) {
  GuideState& g = c->guide;
  hipStream_t s = c->stream;
  HIP_OR_FAIL(c, hipMemsetAsync(g.loss, 0, 4, s));
  {
    const int rc = ensure_wpart(c, guide_train_blocks(n, false, 1));
    if (rc != XNH_OK) return rc;
  }
  HIP_OR_FAIL(c, launch_guide_train(s, g.net, n, dp, dn, dd, dl, lam_scale, nullptr, nullptr, 0, 1, g.grid_g,
                                    g.w_part, g.w_g, g.loss));
  if (apply) {
    g.step++;
    const float b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
    const float bc1 = 1.f - static_cast<float>(std::pow(double(b1), g.step));
    const float bc2 = 1.f - static_cast<float>(std::pow(double(b2), g.step));
    const float lr = g.cfg.learning_rate, gs = 1.f / static_cast<float>(n);
    HIP_OR_FAIL(c, launch_guide_adam(s, g.grid_m, g.grid_h, g.grid_g, g.grid_am, g.grid_av, g.ngrid, lr, b1, b2, eps,
                                     bc1, bc2, gs, nullptr, 0u, 1));
    HIP_OR_FAIL(c, launch_guide_adam(s, g.w_m, g.w_h, g.w_g, g.w_am, g.w_av, size_t(kWTotal), lr, b1, b2, eps, bc1,
                                     bc2, gs, nullptr, 0u, 1));
    HIP_OR_FAIL(c, launch_tile_weights(s, g.w_h, g.w_t));
  }
  return XNH_OK;
}

int xnh_guide_train(xnh_ctx* c, int32_t n, const float* pos, const float* nrm, const float* dir, const float* lam,
                    int32_t apply, float* loss_out, float* grid_grad, float* w_grad) {
  if (!c || !c->guide.on || n <= 0 || !pos || !nrm || !dir || !lam) return XNH_E_INVALID;
  double lsum = 0.0;
  for (int i = 0; i < n; ++i) lsum += lam[i];
  if (!(lsum > 0.0)) return fail(c, XNH_E_INVALID, "training weights sum to zero");
  const float lam_scale = static_cast<float>(double(n) / lsum);
  GuideState& g = c->guide;
  std::vector<void*> tmp;
  float *dp, *dn, *dd, *dl;
  int rc = dev_alloc(c, tmp, &dp, size_t(n) * 3);
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &dn, size_t(n) * 3);
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &dd, size_t(n) * 3);
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &dl, size_t(n));
  if (rc != XNH_OK) {
    free_all(tmp);
    return rc;
  }
  hipStream_t s = c->stream;
  hipError_t e = hipMemcpyAsync(dp, pos, size_t(n) * 12, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dn, nrm, size_t(n) * 12, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dd, dir, size_t(n) * 12, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dl, lam, size_t(n) * 4, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    free_all(tmp);
    return fail(c, XNH_E_HIP, "guide train upload");
  }
  rc = guide_step(c, n, dp, dn, dd, dl, lam_scale, apply != 0);
  float loss = 0.f;
  if (rc == XNH_OK) {
    e = hipMemcpyAsync(&loss, g.loss, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && !apply && grid_grad) e = hipMemcpyAsync(grid_grad, g.grid_g, g.ngrid * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && !apply && w_grad) e = hipMemcpyAsync(w_grad, g.w_g, size_t(kWTotal) * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && !apply) e = hipMemsetAsync(g.grid_g, 0, g.ngrid * 4, s);
    if (e == hipSuccess && !apply) e = hipMemsetAsync(g.w_g, 0, size_t(kWTotal) * 4, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = fail(c, XNH_E_HIP, std::string("guide train: ") + hipGetErrorString(e));
  }
  free_all(tmp);
  if (loss_out) *loss_out = loss / static_cast<float>(n);
  return rc;
}

static int trace_common(xnh_ctx* c, const xnh_ray* rays, int32_t n, bool any, void* out) {
  if (!c || (n > 0 && (!rays || !out)) || n < 0) return XNH_E_INVALID;
  if (!c->has_scene) return fail(c, XNH_E_NO_SCENE, "trace before xnh_upload_scene");
  c->stats = xnh_stats{};
  if (n == 0) return XNH_OK;
  static_assert(sizeof(xnh_ray) == 2 * sizeof(float4), "the kernels read xnh_ray as two float4");
  std::vector<float4> hr(2 * static_cast<size_t>(n));
  std::memcpy(hr.data(), rays, sizeof(xnh_ray) * static_cast<size_t>(n));
  std::vector<void*> tmp;
  float4* d_rays = nullptr;
  void* d_out = nullptr;
  uint32_t* d_spill = nullptr;
  unsigned long long* d_stats = nullptr;
  uint32_t* d_ovf = nullptr;
  const int block = 256;
  const size_t threads = static_cast<size_t>((n + block - 1) / block) * block;
  int rc = dev_upload(c, tmp, &d_rays, hr);
  if (rc == XNH_OK) {
    if (any) rc = dev_alloc(c, tmp, reinterpret_cast<uint8_t**>(&d_out), static_cast<size_t>(n));
    else rc = dev_alloc(c, tmp, reinterpret_cast<float**>(&d_out), static_cast<size_t>(n) * 16);
  }
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &d_spill, static_cast<size_t>(kMaxStack - c->lds_depth) * threads);
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &d_stats, static_cast<size_t>(kNumStat));
  if (rc == XNH_OK) rc = dev_alloc(c, tmp, &d_ovf, 1);
  if (rc != XNH_OK) {
    free_all(tmp);
    return rc;
  }
  hipError_t e = hipMemsetAsync(d_stats, 0, sizeof(unsigned long long) * kNumStat, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(d_ovf, 0, sizeof(uint32_t), c->stream);
  if (e == hipSuccess) e = hipEventRecord(c->ev_begin, c->stream);
  if (e == hipSuccess) {
    e = any ? launch_api_any(c->stream, c->own.S, d_rays, static_cast<uint8_t*>(d_out), n, d_spill, c->lds_depth, d_stats,
                             d_ovf)
            : launch_api_closest(c->stream, c->own.S, d_rays, static_cast<float*>(d_out), n, d_spill, c->lds_depth,
                                 d_stats, d_ovf);
  }
  if (e == hipSuccess) e = hipEventRecord(c->ev_end, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d_out, any ? static_cast<size_t>(n) : static_cast<size_t>(n) * 64, hipMemcpyDeviceToHost,
                       c->stream);
  unsigned long long st[kNumStat] = {};
  uint32_t ovf = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(st, d_stats, sizeof st, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&ovf, d_ovf, sizeof ovf, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev_begin, c->ev_end);
  free_all(tmp);
  if (e != hipSuccess) return fail(c, XNH_E_HIP, std::string("trace: ") + hipGetErrorString(e));
  c->stats.closest_rays = st[kStClosest];
  c->stats.shadow_rays = st[kStShadow];
  c->stats.node_tests = st[kStNodeTests];
  c->stats.interior = st[kStInterior];
  c->stats.tri_tests = st[kStTriTests];
  c->stats.inst_visits = st[kStInstVisits];
  c->stats.shadow_node_tests = st[kStShNodeTests];
  c->stats.shadow_interior = st[kStShInterior];
  c->stats.shadow_tri_tests = st[kStShTriTests];
  c->stats.shadow_inst_visits = st[kStShInstVisits];
  copy_simd_stats(c->stats, st);
  c->stats.render_ms = ms;
  c->stats.stack_overflows = static_cast<int32_t>(ovf);
  if (ovf) return fail(c, XNH_E_OVERFLOW, "traversal stack overflow");
  return XNH_OK;
}

int xnh_trace_closest(xnh_ctx* c, const xnh_ray* rays, int32_t n, xnh_hit* hits) {
  return trace_common(c, rays, n, false, hits);
}
int xnh_trace_any(xnh_ctx* c, const xnh_ray* rays, int32_t n, uint8_t* occ) {
  return trace_common(c, rays, n, true, occ);
}

}  // extern "C"
