// dwave.h — wave-level helpers and batch bookkeeping shared by the
// wavefront kernels (kernels.hip) and the guided-rendering kernels
// (guide.hip): ballot compaction, wave reductions, traversal-counter
// flushes, the per-block view of the current batch's chunk table, and the
// surface record of a closest hit.
#pragma once
#include <hip/hip_runtime.h>

#include "../common/dev_types.h"
#include "dmath.h"
#include "dtrace.h"

namespace xnh {

constexpr float kBounceOffset = 1e-4f;   // math/ray.h:13
constexpr float kShadowOff = 5e-4f;      // math/ray.h:14

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63u); }

// The job table is written by the host before a batch and is read-only while
// kernels run. Read through the constant address space at a wave-uniform
// index, a job's fields are scalar loads (the scalar cache, counted by
// lgkmcnt). As generic loads inside divergent code they were per-lane vector
// loads, each a round trip before the array it names could be read, and each
// wait for one also waited for the wave's pending stores (vmcnt counts both).
// The returned references are generic aliases of the constant-space object;
// once inlined, the compiler infers the address space of the loads through
// them. (A copy by value was kept in scratch memory.) The host compilation
// pass parses these bodies but never runs them; the cast exists only for the
// device.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(4))) const DevJob ConstJob;
__device__ __forceinline__ const DevJob& job_rec(const DevJob* jobs, int job) {
  return reinterpret_cast<const ConstJob*>(reinterpret_cast<uintptr_t>(jobs))[job];
}
#else
__device__ __forceinline__ const DevJob& job_rec(const DevJob* jobs, int job) { return jobs[job]; }
#endif
__device__ __forceinline__ const DevScene& job_scene(const DevJob* jobs, int job) { return job_rec(jobs, job).S; }

// A batch's paths are laid out in 8x8 pixel tiles, so that the 64 primary
// rays of a wave leave one compact patch of the image (a 64x1 row strip
// spreads over eight times the width): the image is cut into bands of 8 rows
// (the last band may be lower), a band into blocks of 8 columns (the last
// block may be narrower), and a block is enumerated row by row. Maps the
// tiled index t in [0, W*H) to the pixel index y*W + x (a bijection).
__device__ __forceinline__ int tiled_pixel(int t, int W, int H) {
  const int band = t / (8 * W);
  const int r = t - band * 8 * W;
  const int hb = min(8, H - 8 * band);
  const int bx = r / (8 * hb);
  const int wb = min(8, W - 8 * bx);
  const int q = r - bx * 8 * hb;
  const int yi = q / wb;
  return (8 * band + yi) * W + 8 * bx + (q - yi * wb);
}

// The camera path behind depth-0 queue entry e of chunk ck: depth 0 holds a
// chunk's paths as its segment 0. Slots are sample-major over the tiled
// pixels (slot = base + sample * npix + t, t the tiled pixel index). The
// entries are in slot order, or, when ck.cam_lg = lg is set, in blocks of
// 64: 2^lg samples of 2^(6-lg) consecutive tiled pixels (lg 3: a row of an
// 8x8 tile; lg 6: one pixel), block after block over the pixels, then over
// the next 2^lg samples, so that a wave's camera rays leave a footprint
// 2^lg times smaller than a tile, while each sample's 2^(6-lg) slots stay
// contiguous (2^lg divides the chunk's samples, 2^(6-lg) its pixels). Its
// seed is the pixel index plus the
// sample index times 0xdeadbeef, its ray the jittered pinhole ray
// (cpu_tiled.cpp:20-29, camera/camera.h:35-37), and rng is the state after
// the two jitter draws. Camera rays are not stored: the depth-0 traversal,
// shading and guide inference each rebuild them here from the entry index,
// bit for bit the same wherever they are made.
struct CameraPath {
  f3 o, d;
  Rng rng;
  uint32_t slot;
};
__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f) {
  return static_cast<uint32_t>((static_cast<uint64_t>(n) * f.m) >> f.p);
}
__device__ __forceinline__ CameraPath camera_path(const DevChunk& ck, const DevJob& J, uint32_t e) {
  // the divisions by the chunk's and the job's sizes are multiply-shifts
  // (FastDiv); the rest are shifts except in the last tile row and column
  const uint32_t local = e - static_cast<uint32_t>(ck.qbase);
  const uint32_t npix = static_cast<uint32_t>(ck.npix);
  uint32_t s, t;
  if (ck.cam_lg) {
    const uint32_t lg = static_cast<uint32_t>(ck.cam_lg), lp = 6u - lg;   // log2 of the samples, pixels per block
    const uint32_t blk = local >> 6, l = local & 63u;
    const uint32_t sg = fdiv(blk, ck.sdiv);   // / (npix >> lp)
    s = (sg << lg) + (l >> lp);
    t = ((blk - sg * (npix >> lp)) << lp) + (l & ((1u << lp) - 1u));
  } else {
    s = fdiv(local, ck.sdiv);   // / npix
    t = local - s * npix;
  }
  // tiled index -> pixel (x, y): the steps of tiled_pixel
  const int W = J.W, H = J.H;
  const int ti = static_cast<int>(t);
  const int band = static_cast<int>(fdiv(t, J.band));   // / (8 W)
  const int r = ti - band * 8 * W;
  const int hb = min(8, H - 8 * band);
  const int bx = hb == 8 ? r >> 6 : r / (8 * hb);
  const int wb = min(8, W - 8 * bx);
  const int q = r - bx * 8 * hb;
  const int yi = wb == 8 ? q >> 3 : q / wb;
  const int x = 8 * bx + (q - yi * wb), y = 8 * band + yi;
  const int pixel = y * W + x;
  const int spp = ck.spp_first + static_cast<int>(s);
  const uint64_t seed = static_cast<uint64_t>(pixel) + static_cast<uint64_t>(spp) * 0xdeadbeefULL;
  CameraPath p;
  p.rng = rng_for(seed);
  const float fx = static_cast<float>(x) + p.rng.f();
  const float fy = static_cast<float>(y) + p.rng.f();
  const float u = fx / static_cast<float>(J.W), w = fy / static_cast<float>(J.H);
  p.o = f3{J.cam.eye[0], J.cam.eye[1], J.cam.eye[2]};
  const f3 llc{J.cam.llc[0], J.cam.llc[1], J.cam.llc[2]};
  const f3 hz{J.cam.horiz[0], J.cam.horiz[1], J.cam.horiz[2]};
  const f3 vt{J.cam.vert[0], J.cam.vert[1], J.cam.vert[2]};
  p.d = normalize(llc + u * hz + w * vt - p.o);
  p.slot = static_cast<uint32_t>(ck.base) + s * npix + t;
  return p;
}

// Wave-aggregated append: one atomic per wave, ranks by ballot prefix.
// Must be reached by every lane of the wave.
__device__ __forceinline__ uint32_t wave_append(uint32_t* ctr, bool want) {
  const unsigned long long mask = __ballot(want);
  if (mask == 0ull) return 0u;
  const int lane = lane_id();
  const uint32_t rank = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
  const int leader = __ffsll(static_cast<long long>(mask)) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(ctr, static_cast<uint32_t>(__popcll(mask)));
  base = __shfl(base, leader);
  return base + rank;
}

// Two appends to one counter with one atomic: the `a` lanes take the first
// popcount(a) places of the wave's block, the `b` lanes the places after
// them. Must be reached by every lane of the wave.
__device__ __forceinline__ void wave_append2(uint32_t* ctr, bool a, bool b, uint32_t& ia, uint32_t& ib) {
  const unsigned long long ma = __ballot(a);
  const unsigned long long mb = __ballot(b);
  ia = ib = 0u;
  if ((ma | mb) == 0ull) return;
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t na = static_cast<uint32_t>(__popcll(ma));
  const int leader = __ffsll(static_cast<long long>(ma | mb)) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(ctr, na + static_cast<uint32_t>(__popcll(mb)));
  base = __shfl(base, leader);
  ia = base + static_cast<uint32_t>(__popcll(ma & below));
  ib = base + na + static_cast<uint32_t>(__popcll(mb & below));
}

// The shading kernel's appends of one pass: shadow records (the `a` lanes,
// then the `b` lanes) to ctr_s and continuations (`c` lanes) to ctr_c. The
// leader issues both atomics before either result is read, so their round
// trips overlap (one after the other, the second atomic's wait had cost as
// much as the rest of the pass). Must be reached by every lane of the wave.
__device__ __forceinline__ void wave_append_pass(uint32_t* ctr_s, bool a, bool b, uint32_t& ia, uint32_t& ib,
                                                 uint32_t* ctr_c, bool c, uint32_t& ic) {
  const unsigned long long ma = __ballot(a), mb = __ballot(b), mc = __ballot(c);
  ia = ib = ic = 0u;
  if ((ma | mb | mc) == 0ull) return;
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t na = static_cast<uint32_t>(__popcll(ma));
  const uint32_t ns = na + static_cast<uint32_t>(__popcll(mb)), nc = static_cast<uint32_t>(__popcll(mc));
  const int leader = __ffsll(static_cast<long long>(ma | mb | mc)) - 1;
  uint32_t bs = 0u, bc = 0u;
  if (lane == leader) {
    if (ns) bs = atomicAdd(ctr_s, ns);
    if (nc) bc = atomicAdd(ctr_c, nc);
  }
  // one broadcast of both results: the wait for either comes after both
  // atomics are issued
  const unsigned long long both =
      __shfl(static_cast<unsigned long long>(bc) << 32 | static_cast<unsigned long long>(bs), leader);
  bs = static_cast<uint32_t>(both);
  bc = static_cast<uint32_t>(both >> 32);
  ia = bs + static_cast<uint32_t>(__popcll(ma & below));
  ib = bs + na + static_cast<uint32_t>(__popcll(mb & below));
  ic = bc + static_cast<uint32_t>(__popcll(mc & below));
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Wave-reduced ray and (with `work`) traversal counters; `any` selects the
// shadow slots.
// The persistent traversal kernels' counts, summed over the work-group and
// added with one atomic per counter and work-group: the waves of a launch
// finish together, and one same-address atomic per wave at the end made a
// serial tail. Every thread of the work-group must reach it (256 threads).
__device__ __forceinline__ void flush_counts_block(unsigned long long* stats, const TravCount& c, uint32_t rays,
                                                   bool any, bool work) {
  constexpr int kN = 10;
  __shared__ unsigned long long red[kN][256 / 64];
  const int w = static_cast<int>(threadIdx.x >> 6), nw = static_cast<int>(blockDim.x >> 6);
  const unsigned long long v[kN] = {
      wave_sum(rays),
      work ? wave_sum(c.node_tests) : 0ull,
      work ? wave_sum(c.interior) : 0ull,
      work ? wave_sum(c.tri_tests) : 0ull,
      work ? wave_sum(c.inst_visits) : 0ull,
      work ? wave_sum(c.int_iters) : 0ull,
      work ? wave_sum(c.leaf_iters) : 0ull,
      work ? wave_sum(c.leaf_steps) : 0ull,
      work ? wave_sum(c.round_iters) : 0ull,
      work ? wave_sum(c.round_lanes) : 0ull};
  if (lane_id() == 0)
    for (int k = 0; k < kN; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < static_cast<unsigned>(kN)) {
    unsigned long long t = 0;
    for (int k = 0; k < nw; ++k) t += red[threadIdx.x][k];
    const int k = static_cast<int>(threadIdx.x);
    int slot;
    if (k == 0) slot = any ? kStShadow : kStClosest;
    else if (k < 5) slot = (any ? kStShNodeTests : kStNodeTests) + k - 1;
    else slot = (any ? kStShIntIters : kStIntIters) + k - 5;
    if (t) atomicAdd(&stats[slot], t);
  }
}

__device__ __forceinline__ void flush_counts(unsigned long long* stats, const TravCount& c, uint32_t rays, bool any,
                                             bool work = true) {
  const unsigned long long r = wave_sum(rays);
  if (lane_id() == 0 && r) atomicAdd(&stats[any ? kStShadow : kStClosest], r);
  if (!work) return;
  const unsigned long long a = wave_sum(c.node_tests), b = wave_sum(c.interior), t = wave_sum(c.tri_tests),
                           i = wave_sum(c.inst_visits);
  const int base = any ? kStShNodeTests : kStNodeTests;
  if (lane_id() == 0) {
    if (a) atomicAdd(&stats[base], a);
    if (b) atomicAdd(&stats[base + 1], b);
    if (t) atomicAdd(&stats[base + 2], t);
    if (i) atomicAdd(&stats[base + 3], i);
  }
  const unsigned long long p[5] = {wave_sum(c.int_iters), wave_sum(c.leaf_iters), wave_sum(c.leaf_steps),
                                   wave_sum(c.round_iters), wave_sum(c.round_lanes)};
  const int pbase = any ? kStShIntIters : kStIntIters;
  if (lane_id() == 0)
    for (int k = 0; k < 5; ++k)
      if (p[k]) atomicAdd(&stats[pbase + k], p[k]);
}

__device__ __forceinline__ TravStack make_stack(uint32_t* spill, int lds_depth) {
  TravStack st;
  st.lds_col = static_cast<int>(threadIdx.x);
  st.lds_stride = static_cast<int>(blockDim.x);
  st.spill = spill + (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x);
  st.spill_stride = static_cast<int>(gridDim.x * blockDim.x);
  st.lds_depth = lds_depth;
  st.reset();
  return st;
}

// Surface data of a closest hit: the reference's HitRecord after
// intersect_triangle (geometry/mesh.h:104-129) and TLAS::fixup_hit
// (geometry/tlas.cpp:156-168).
struct Surface {
  f3 pos, n, gn;
  int mat;
  uint32_t tri;
  int32_t inst_id;
  int32_t lmap;
  bool front;
};

__device__ __forceinline__ f3 normal_out(const InstRec& in, f3 n) {
  return normalize(f3{in.c0.x * n.x + in.c0.y * n.y + in.c0.z * n.z, in.c1.x * n.x + in.c1.y * n.y + in.c1.z * n.z,
                      in.c2.x * n.x + in.c2.y * n.y + in.c2.z * n.z});
}

// A primitive hit: the primitive's own normal (facing the ray) and uv in
// instance space, then TLAS::fixup_hit (tlas.cpp:156-168).
__device__ __forceinline__ Surface resolve_prim(const DevScene& S, const HitRef& h, const InstRec& in, f3 o, f3 d,
                                               float tmin, float& u, float& v) {
  const uint32_t pi = h.slot & ~kPrimHit;
  const PrimRec pr = load_prim(S.prims, pi);
  const f3 lo = in.c0 * o.x + in.c1 * o.y + in.c2 * o.z + in.c3;
  const f3 ld = in.c0 * d.x + in.c1 * d.y + in.c2 * d.z;
  const f3 nl = prim_normal(pr, lo, ld, tmin, h.t, u, v);
  Surface s;
  s.tri = pi;
  s.pos = o + h.t * d;
  s.n = normal_out(in, nl);
  s.gn = s.n;
  s.front = dot(s.gn, d) < 0.f;
  s.mat = pr.mat;
  s.inst_id = in.id;
  s.lmap = -1;
  return s;
}

__device__ __forceinline__ Surface resolve(const DevScene& S, const HitRef& h, f3 o, f3 d, float tmin = 0.f) {
  const InstRec in = load_inst(S.inst, h.inst);
  if (h.slot & kPrimHit) {
    float u, v;
    return resolve_prim(S, h, in, o, d, tmin, u, v);
  }
  const float4 a = S.tris[3 * h.slot], b = S.tris[3 * h.slot + 1], c = S.tris[3 * h.slot + 2];
  const f3 e1{a.w, b.x, b.y}, e2{b.z, b.w, c.x};
  Surface s;
  s.tri = __float_as_uint(c.y);
  const uint32_t gtri = (in.tri_base & ~kFlatBit) + s.tri;
  const f3 gl = normalize(cross(e1, e2));
  f3 nl = gl;
  if (!(in.tri_base & kFlatBit)) {
    const int i0 = S.vidx[3 * gtri], i1 = S.vidx[3 * gtri + 1], i2 = S.vidx[3 * gtri + 2];
    const f3 n0 = xyz(S.vnrm[i0]), n1 = xyz(S.vnrm[i1]), n2 = xyz(S.vnrm[i2]);
    nl = normalize(n0 * (1.f - h.u - h.v) + n1 * h.u + n2 * h.v);
  }
  s.pos = o + h.t * d;
  s.n = normal_out(in, nl);
  s.gn = normal_out(in, gl);
  s.front = dot(s.gn, d) < 0.f;
  s.mat = S.tmat[gtri];
  s.inst_id = in.id;
  s.lmap = in.lmap;
  return s;
}

// The part of `resolve` the guide needs: position, world geometric normal,
// material and the light-map lookup of the hit triangle.
struct GeoSurface {
  f3 pos, gn;
  int mat;
  uint32_t tri;
  int32_t lmap;
};

__device__ __forceinline__ GeoSurface resolve_geo(const DevScene& S, const HitRef& h, f3 o, f3 d, float tmin = 0.f) {
  const InstRec in = load_inst(S.inst, h.inst);
  if (h.slot & kPrimHit) {
    float u, v;
    const Surface f = resolve_prim(S, h, in, o, d, tmin, u, v);
    return GeoSurface{f.pos, f.gn, f.mat, f.tri, -1};
  }
  const float4 a = S.tris[3 * h.slot], b = S.tris[3 * h.slot + 1], c = S.tris[3 * h.slot + 2];
  const f3 e1{a.w, b.x, b.y}, e2{b.z, b.w, c.x};
  GeoSurface s;
  s.tri = __float_as_uint(c.y);
  s.pos = o + h.t * d;
  s.gn = normal_out(in, normalize(cross(e1, e2)));
  s.mat = S.tmat[(in.tri_base & ~kFlatBit) + s.tri];
  s.lmap = in.lmap;
  return s;
}

// Gradient sky (scene/environment.h:34-46).
__device__ __forceinline__ f3 sky_eval(const DevSky& k, f3 d) {
  const f3 hz{k.horizon[0], k.horizon[1], k.horizon[2]};
  if (d.y < 0.f) {
    const float t = smin(-d.y, 1.f);
    return k.intensity * lerp3(hz, f3{k.ground[0], k.ground[1], k.ground[2]}, t);
  }
  const float t = pow_r(d.y, 1.0f / k.sharpness);
  return k.intensity * lerp3(hz, f3{k.zenith[0], k.zenith[1], k.zenith[2]}, t);
}


// ---- batch bookkeeping ------------------------------------------------------

// Per-block copy of the current batch's chunk table with the entry count of
// every queue segment and prefix sums of 64-entry work units over chunks
// (traversal / shading) or of pixels (accumulation).
struct BatchView {
  int batch, nch;
  DevChunk ch[kMaxChunks];
  uint32_t cnt[kMaxChunks][kQSeg];
  uint32_t prefix[kMaxChunks + 1];
};

// Loads the batch and builds prefix[] from the segment counts. counts: the
// queue's counter of chunk 0, segment 0, as ctr_path / ctr_shadow give it,
// with mul 1 (path queue) or 2 (shadow queue); null: prefix over pixels.
// A count is clamped to its segment (segment 0 of a path queue may hold the
// chunk's whole region: depth 0), so reads stay inside the chunk's region
// even if an append ever overflowed a segment. Block-wide; ends with a
// barrier.
__device__ __forceinline__ void load_batch(BatchView& v, const WaveBuffers& B, const uint32_t* counts,
                                           bool pixels, uint32_t mul = 1u) {
  if (threadIdx.x == 0) {
    v.batch = static_cast<int>(B.counters[kBatch]);
    v.nch = B.nchunks[v.batch];
  }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < v.nch) v.ch[threadIdx.x] = B.chunks[v.batch * kMaxChunks + threadIdx.x];
  __syncthreads();
  if (counts)
    for (int t = static_cast<int>(threadIdx.x); t < v.nch * kQSeg; t += static_cast<int>(blockDim.x)) {
      const int ci = t / kQSeg, x = t % kQSeg;
      const uint32_t qseg = static_cast<uint32_t>(v.ch[ci].qseg);
      const uint32_t room = mul * ((mul == 1u && x == 0) ? static_cast<uint32_t>(kQSeg) * qseg : qseg);
      v.cnt[ci][x] = min(counts[t * kCtrStride], room);
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    v.prefix[0] = 0;
    for (int c = 0; c < v.nch; ++c) {
      if (pixels) {
        acc += static_cast<uint32_t>(v.ch[c].npix);
      } else {
        for (int x = 0; x < kQSeg; ++x) acc += (v.cnt[c][x] + 63u) >> 6;
      }
      v.prefix[c + 1] = acc;
    }
  }
  __syncthreads();
}

// Last chunk c with prefix[c] <= u: the (non-empty) chunk holding item u.
__device__ __forceinline__ int find_chunk(const BatchView& v, uint32_t u) {
  int lo = 0, hi = v.nch;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (v.prefix[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Unit uc of chunk ci's units (segment by segment): returns the queue
// position of its segment's first entry (mul = 1 path queue, 2 shadow
// queue) and sets its first entry's offset in that segment and the
// segment's entry count.
__device__ __forceinline__ uint32_t unit_segment(const BatchView& v, int ci, uint32_t uc, uint32_t mul, uint32_t& first,
                                                 uint32_t& lim) {
  uint32_t acc = 0;
  int x = 0;
  for (; x < kQSeg - 1; ++x) {
    const uint32_t n = (v.cnt[ci][x] + 63u) >> 6;
    if (uc < acc + n) break;
    acc += n;
  }
  first = (uc - acc) * 64u;
  lim = v.cnt[ci][x];
  const DevChunk& ck = v.ch[ci];
  return mul * (static_cast<uint32_t>(ck.qbase) + static_cast<uint32_t>(x) * static_cast<uint32_t>(ck.qseg));
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }


}  // namespace xnh
