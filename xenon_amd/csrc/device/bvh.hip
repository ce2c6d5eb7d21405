// bvh.hip — BLAS construction and refit on the device.
//
// Build: a linear BVH (Lauterbach et al. 2009; the parallel radix-tree
// construction of Karras 2012). Triangle box centres are quantised to 10
// bits per axis and interleaved into 30-bit Morton codes, sorted with their
// triangle indices (rocPRIM's radix sort; ties keep index order), and every
// internal node of the binary radix tree over the sorted codes is found
// independently from the longest common prefixes of its neighbours (equal
// codes are told apart by their positions). Subtree sizes come bottom-up,
// then every node's place in the output is computed from its ancestors so
// that the tree has the reference's node format (blas.h:21-28): 32-byte
// nodes, the two children of a node adjacent (right = left + 1) and after
// it, pairs starting at odd indices (64-byte aligned after the upload's
// odd mesh offset), one triangle per leaf.
//
// Refit: boxes of any tree in that format for moved vertices: each leaf
// takes the bounds of its triangles, and each interior node takes the union
// of its children once both are done (the second child to arrive computes
// it, ordered by an atomic counter per node and agent-scope fences, since
// the eight XCDs' L2 caches are not coherent with each other).
//
// Every walk is bounded and every computed index is checked before it is
// used; a violation sets the error word instead of touching memory.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "launch.h"

namespace xnh {
namespace {

constexpr int kT = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kMaxWalk = 4096;   // bound of any upward walk (radix trees are at most 64 deep)

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ void tri_bounds(const float* vx, const float* vy, const float* vz, const int32_t* idx,
                                           uint32_t t, float* lo, float* hi) {
  const int32_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
  const float x[3] = {vx[a], vx[b], vx[c]}, y[3] = {vy[a], vy[b], vy[c]}, z[3] = {vz[a], vz[b], vz[c]};
  lo[0] = fminf(fminf(x[0], x[1]), x[2]);
  lo[1] = fminf(fminf(y[0], y[1]), y[2]);
  lo[2] = fminf(fminf(z[0], z[1]), z[2]);
  hi[0] = fmaxf(fmaxf(x[0], x[1]), x[2]);
  hi[1] = fmaxf(fmaxf(y[0], y[1]), y[2]);
  hi[2] = fmaxf(fmaxf(z[0], z[1]), z[2]);
}

// Triangle box centres and per-block bounds of them.
__global__ void __launch_bounds__(kT) k_lb_centres(const float* vx, const float* vy, const float* vz,
                                                   const int32_t* idx, int n, float* cen, float* pb) {
  __shared__ float red[6][kT];
  float mn[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
  float mx[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
  for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < n;
       i += static_cast<int>(gridDim.x * blockDim.x)) {
    float lo[3], hi[3];
    tri_bounds(vx, vy, vz, idx, static_cast<uint32_t>(i), lo, hi);
    for (int k = 0; k < 3; ++k) {
      const float c = 0.5f * (lo[k] + hi[k]);
      cen[3 * static_cast<size_t>(i) + k] = c;
      mn[k] = fminf(mn[k], c);
      mx[k] = fmaxf(mx[k], c);
    }
  }
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = mn[k];
    red[3 + k][threadIdx.x] = mx[k];
  }
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) pb[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// The centres' bounds from the per-block bounds (one block).
__global__ void __launch_bounds__(64) k_lb_bounds(const float* pb, int nb, float* bounds) {
  if (threadIdx.x < 6) {
    const bool lo = threadIdx.x < 3;
    float v = lo ? __builtin_huge_valf() : -__builtin_huge_valf();
    for (int k = 0; k < nb; ++k) v = lo ? fminf(v, pb[6 * k + threadIdx.x]) : fmaxf(v, pb[6 * k + threadIdx.x]);
    bounds[threadIdx.x] = v;
  }
}

// Morton code of each centre within the centres' bounds.
__global__ void __launch_bounds__(kT) k_lb_morton(const float* cen, int n, const float* bounds, uint32_t* keys,
                                                  uint32_t* vals) {
  const float b[6] = {bounds[0], bounds[1], bounds[2], bounds[3], bounds[4], bounds[5]};
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const float ext = b[3 + k] - b[k];
    const float u = ext > 0.f ? (cen[3 * static_cast<size_t>(i) + k] - b[k]) / ext : 0.f;
    q[k] = static_cast<uint32_t>(fminf(fmaxf(u * 1024.f, 0.f), 1023.f));
  }
  keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  vals[i] = static_cast<uint32_t>(i);
}

// Length of the common prefix of sorted keys a and b (positions break ties);
// -1 outside the array.
__device__ __forceinline__ int lcp(const uint32_t* keys, int n, int a, long long b) {
  if (b < 0 || b >= n) return -1;
  const uint32_t ka = keys[a], kb = keys[static_cast<int>(b)];
  if (ka != kb) return __clz(static_cast<int>(ka ^ kb));
  return 32 + __clz(static_cast<int>(static_cast<uint32_t>(a) ^ static_cast<uint32_t>(b)));
}

// Internal node i of the radix tree (Karras 2012, fig. 4): its key range
// [i, j] or [j, i], the split gamma, the two children. Node ids: internal
// 0..n-2 (0 is the root), leaf k is n-1+k.
__global__ void __launch_bounds__(kT) k_lb_radix(const uint32_t* keys, int n, uint32_t* left, uint32_t* right,
                                                 uint32_t* parent, uint32_t* err) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n - 1) return;
  const int d = lcp(keys, n, i, i + 1) - lcp(keys, n, i, i - 1) >= 0 ? 1 : -1;
  const int dmin = lcp(keys, n, i, i - d);
  long long lmax = 2;
  while (lcp(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  long long l = 0;
  for (long long t = lmax >> 1; t >= 1; t >>= 1)
    if (lcp(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const long long j = i + l * d;
  const int dnode = lcp(keys, n, i, j);
  long long s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (lcp(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const long long gamma = i + s * d + (d < 0 ? -1 : 0);
  const long long lo = i < j ? i : j, hi = i < j ? j : i;
  if (gamma < 0 || gamma + 1 >= n || gamma < lo || gamma + 1 > hi) {
    atomicOr(err, 1u);
    return;
  }
  const uint32_t a = lo == gamma ? static_cast<uint32_t>(n - 1 + gamma) : static_cast<uint32_t>(gamma);
  const uint32_t b = hi == gamma + 1 ? static_cast<uint32_t>(n + gamma) : static_cast<uint32_t>(gamma + 1);
  left[i] = a;
  right[i] = b;
  parent[a] = static_cast<uint32_t>(i);
  parent[b] = static_cast<uint32_t>(i);
}

// Subtree sizes, bottom-up: the second child to reach a node sums it.
__global__ void __launch_bounds__(kT) k_lb_sizes(int n, const uint32_t* left, const uint32_t* right,
                                                 const uint32_t* parent, uint32_t* size, uint32_t* flag,
                                                 uint32_t* err) {
  const int k = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= n) return;
  const uint32_t nn = static_cast<uint32_t>(2 * n - 1);
  uint32_t x = static_cast<uint32_t>(n - 1 + k);
  size[x] = 1u;
  for (int step = 0; step < kMaxWalk; ++step) {
    const uint32_t p = parent[x];
    if (p >= static_cast<uint32_t>(n - 1)) {
      if (x != 0u) atomicOr(err, 2u);
      return;
    }
    __threadfence();
    if (atomicAdd(&flag[p], 1u) == 0u) return;
    __threadfence();
    const uint32_t a = left[p], b = right[p];
    if (a >= nn || b >= nn) {
      atomicOr(err, 2u);
      return;
    }
    size[p] = 1u + __hip_atomic_load(&size[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) +
              __hip_atomic_load(&size[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p == 0u) return;
    x = p;
  }
  atomicOr(err, 4u);
}

// Output index of every node: root 0; a node's children pair opens the
// region of its descendants, which starts right after its parent's pair
// for a left child and after the left sibling's descendants for a right one.
__global__ void __launch_bounds__(kT) k_lb_place(int n, const uint32_t* left, const uint32_t* right,
                                                 const uint32_t* parent, const uint32_t* size, uint32_t* pos,
                                                 uint32_t* err) {
  const int x = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int nn = 2 * n - 1;
  if (x >= nn) return;
  if (x == 0) {
    pos[0] = 0u;
    return;
  }
  uint32_t p = parent[x];
  if (p >= static_cast<uint32_t>(n - 1)) {
    atomicOr(err, 8u);
    return;
  }
  uint64_t acc = 1u + (right[p] == static_cast<uint32_t>(x) ? 1u : 0u);
  uint32_t a = p;
  int step = 0;
  while (a != 0u && step++ < kMaxWalk) {
    const uint32_t pa = parent[a];
    if (pa >= static_cast<uint32_t>(n - 1)) {
      atomicOr(err, 8u);
      return;
    }
    acc += 2u + (right[pa] == a ? size[left[pa]] - 1u : 0u);
    a = pa;
  }
  if (a != 0u || acc >= static_cast<uint64_t>(nn)) {
    atomicOr(err, 8u);
    return;
  }
  pos[x] = static_cast<uint32_t>(acc);
}

// Nodes in the output order: interior nodes name their first child, leaves
// their slot (one triangle each); slot k holds the k-th triangle in Morton
// order. Boxes are filled by the refit.
__global__ void __launch_bounds__(kT) k_lb_emit(int n, const uint32_t* left, const uint32_t* right,
                                                const uint32_t* pos, const uint32_t* sorted_tri,
                                                xnh_bvh_node* nodes, uint32_t* prims, uint32_t* err) {
  const int x = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const int nn = 2 * n - 1;
  if (x >= nn) return;
  const uint32_t q = pos[x];
  if (q >= static_cast<uint32_t>(nn)) {
    atomicOr(err, 16u);
    return;
  }
  xnh_bvh_node o{};
  if (x < n - 1) {
    const uint32_t a = pos[left[x]], b = pos[right[x]];
    if (b != a + 1u || a <= q) {
      atomicOr(err, 16u);
      return;
    }
    o.child = a;
    o.count = 0u;
  } else {
    const uint32_t k = static_cast<uint32_t>(x - (n - 1));
    o.child = k;
    o.count = 1u;
    prims[k] = sorted_tri[k];
  }
  nodes[q] = o;
}

__global__ void __launch_bounds__(kT) k_rf_parents(const xnh_bvh_node* nodes, int nn, uint32_t* parent,
                                                   uint32_t* err) {
  const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= nn) return;
  if (j == 0) parent[0] = kNone;
  const xnh_bvh_node& v = nodes[j];
  if (v.count == 0u) {
    // children follow their parent, so every upward walk strictly descends
    if (v.child <= static_cast<uint32_t>(j) || v.child >= static_cast<uint32_t>(nn - 1)) {
      atomicOr(err, 128u);
      return;
    }
    parent[v.child] = static_cast<uint32_t>(j);
    parent[v.child + 1] = static_cast<uint32_t>(j);
  }
}

// Leaf boxes from the vertices, then unions up to the root.
__global__ void __launch_bounds__(kT) k_rf_boxes(xnh_bvh_node* nodes, int nn, const uint32_t* prims, int nprims,
                                                 const float* vx, const float* vy, const float* vz,
                                                 const int32_t* idx, const uint32_t* parent, uint32_t* flag,
                                                 uint32_t* err) {
  const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= nn) return;
  xnh_bvh_node* v = &nodes[j];
  if (v->count == 0u) return;
  if (static_cast<uint64_t>(v->child) + v->count > static_cast<uint64_t>(nprims)) {
    atomicOr(err, 256u);
    return;
  }
  float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
  float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
  for (uint32_t s = v->child; s < v->child + v->count; ++s) {   // the host checked the slot range
    float a[3], b[3];
    tri_bounds(vx, vy, vz, idx, prims[s], a, b);
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(lo[k], a[k]);
      hi[k] = fmaxf(hi[k], b[k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    v->lo[k] = lo[k];
    v->hi[k] = hi[k];
  }
  uint32_t x = static_cast<uint32_t>(j);
  for (int step = 0; step < kMaxWalk; ++step) {
    const uint32_t p = parent[x];
    if (p == kNone) return;
    if (p >= static_cast<uint32_t>(nn)) {
      atomicOr(err, 32u);
      return;
    }
    __threadfence();
    if (atomicAdd(&flag[p], 1u) == 0u) return;
    __threadfence();
    xnh_bvh_node* P = &nodes[p];
    const uint32_t c = P->child;
    if (P->count != 0u || c + 1u >= static_cast<uint32_t>(nn)) {
      atomicOr(err, 32u);
      return;
    }
    const float* l0 = nodes[c].lo;
    const float* h0 = nodes[c].hi;
    const float* l1 = nodes[c + 1].lo;
    const float* h1 = nodes[c + 1].hi;
    for (int k = 0; k < 3; ++k) {
      const float a0 = __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(l0 + k), __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT));
      const float a1 = __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(l1 + k), __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT));
      const float b0 = __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(h0 + k), __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT));
      const float b1 = __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(h1 + k), __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT));
      P->lo[k] = fminf(a0, a1);
      P->hi[k] = fmaxf(b0, b1);
    }
    x = p;
  }
  atomicOr(err, 64u);
}

int blocks_for(long long n) { return static_cast<int>((n + kT - 1) / kT); }

}  // namespace

hipError_t refit_blas_device(hipStream_t s, xnh_bvh_node* nodes, int nn, const uint32_t* prims, int nprims,
                             const float* vx, const float* vy, const float* vz, const int32_t* idx, uint32_t* parent,
                             uint32_t* flag, uint32_t* err) {
  if (nn <= 0) return hipSuccess;
  hipError_t e = hipMemsetAsync(flag, 0, sizeof(uint32_t) * nn, s);
  if (e == hipSuccess) e = hipMemsetAsync(parent, 0xFF, sizeof(uint32_t) * nn, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_rf_parents, dim3(blocks_for(nn)), dim3(kT), 0, s, nodes, nn, parent, err);
  hipLaunchKernelGGL(k_rf_boxes, dim3(blocks_for(nn)), dim3(kT), 0, s, nodes, nn, prims, nprims, vx, vy, vz, idx,
                     parent, flag, err);
  return hipGetLastError();
}

hipError_t build_blas_device(hipStream_t s, const float* vx, const float* vy, const float* vz, const int32_t* idx,
                             int n, xnh_bvh_node* nodes, uint32_t* prims, uint32_t* err) {
  if (n < 2) return hipErrorInvalidValue;   // the caller builds single-triangle trees itself
  const int nn = 2 * n - 1;
  const int nb = blocks_for(n) < 1024 ? blocks_for(n) : 1024;
  std::vector<void*> mem;
  hipError_t e = hipSuccess;
  auto alloc = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (e == hipSuccess) e = hipMalloc(&p, bytes < 4 ? 4 : bytes);
    if (e == hipSuccess) mem.push_back(p);
    return p;
  };
  float* cen = static_cast<float*>(alloc(sizeof(float) * 3 * n));
  float* pb = static_cast<float*>(alloc(sizeof(float) * 6 * nb));
  float* bounds = static_cast<float*>(alloc(sizeof(float) * 6));
  uint32_t* k0 = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * n));
  uint32_t* k1 = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * n));
  uint32_t* v0 = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * n));
  uint32_t* v1 = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * n));
  uint32_t* left = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * n));
  uint32_t* right = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * n));
  uint32_t* parent = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * nn));
  uint32_t* size = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * nn));
  uint32_t* flag = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * nn));
  uint32_t* pos = static_cast<uint32_t*>(alloc(sizeof(uint32_t) * nn));
  size_t tmp_bytes = 0;
  if (e == hipSuccess) e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, k0, k1, v0, v1, static_cast<size_t>(n), 0, 30, s);
  void* tmp = alloc(tmp_bytes);
  if (e == hipSuccess) e = hipMemsetAsync(parent, 0xFF, sizeof(uint32_t) * nn, s);
  if (e == hipSuccess) e = hipMemsetAsync(flag, 0, sizeof(uint32_t) * nn, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_lb_centres, dim3(nb), dim3(kT), 0, s, vx, vy, vz, idx, n, cen, pb);
    hipLaunchKernelGGL(k_lb_bounds, dim3(1), dim3(64), 0, s, pb, nb, bounds);
    hipLaunchKernelGGL(k_lb_morton, dim3(blocks_for(n)), dim3(kT), 0, s, cen, n, bounds, k0, v0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = rocprim::radix_sort_pairs(tmp, tmp_bytes, k0, k1, v0, v1, static_cast<size_t>(n), 0, 30, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_lb_radix, dim3(blocks_for(n - 1)), dim3(kT), 0, s, k1, n, left, right, parent, err);
    hipLaunchKernelGGL(k_lb_sizes, dim3(blocks_for(n)), dim3(kT), 0, s, n, left, right, parent, size, flag, err);
    hipLaunchKernelGGL(k_lb_place, dim3(blocks_for(nn)), dim3(kT), 0, s, n, left, right, parent, size, pos, err);
    hipLaunchKernelGGL(k_lb_emit, dim3(blocks_for(nn)), dim3(kT), 0, s, n, left, right, pos, v1, nodes, prims, err);
    e = hipGetLastError();
  }
  // boxes: the generic refit over the emitted layout (parent and flag reused)
  if (e == hipSuccess) e = refit_blas_device(s, nodes, nn, prims, n, vx, vy, vz, idx, parent, flag, err);
  const hipError_t es = hipStreamSynchronize(s);
  if (e == hipSuccess) e = es;
  for (void* p : mem) (void)hipFree(p);
  return e;
}

}  // namespace xnh
