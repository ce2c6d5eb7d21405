// kernels.hip — wavefront path-tracing kernels for gfx950 (wave64).
//
// One batch renders `nsamples` samples of every pixel, C = pixels*nsamples
// paths in flight, through this stream-ordered sequence (no host reads
// between kernels; every work count lives in device memory):
//
//   raygen ─► [ closest(d) ─► shade(d) ─► shadow(d) ]  d = 0 .. max_bounces-1
//          ─► accumulate
//
// closest / shadow are persistent: a grid sized to fill the chip, each wave
// fetching 64 queue entries at a time from a device cursor, so divergent
// traversal work is balanced across waves. shade appends survivors and
// shadow rays with one atomic per wave (ballot + popcount compaction), so
// the next bounce's queue is dense and every queue access is coalesced.
//
// Integrator semantics: render/cpu_tiled.cpp:13-237 (see the per-stage
// comments); BSDF: dbsdf.h; traversal: dtrace.h.
#include <hip/hip_runtime.h>

#include "../common/dev_types.h"
#include "dbsdf.h"
#include "dmath.h"
#include "dtrace.h"
#include "launch.h"

namespace xnh {

namespace {

constexpr float kBounceOffset = 1e-4f;   // math/ray.h:13
constexpr float kShadowOff = 5e-4f;      // math/ray.h:14

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63u); }

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

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Wave-reduced traversal counters; `any` selects the shadow slots.
__device__ __forceinline__ void flush_counts(unsigned long long* stats, const TravCount& c, uint32_t rays, bool any) {
  const unsigned long long a = wave_sum(c.node_tests), b = wave_sum(c.interior), t = wave_sum(c.tri_tests),
                           i = wave_sum(c.inst_visits);
  const unsigned long long r = wave_sum(rays);
  const int base = any ? kStShNodeTests : kStNodeTests;
  if (lane_id() == 0) {
    if (a) atomicAdd(&stats[base], a);
    if (b) atomicAdd(&stats[base + 1], b);
    if (t) atomicAdd(&stats[base + 2], t);
    if (i) atomicAdd(&stats[base + 3], i);
    if (r) atomicAdd(&stats[any ? kStShadow : kStClosest], r);
  }
}

__device__ __forceinline__ TravStack make_stack(uint32_t* lds, uint32_t* spill, int lds_depth) {
  TravStack st;
  st.lds = lds + threadIdx.x;
  st.lds_stride = static_cast<int>(blockDim.x);
  st.spill = spill + (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x);
  st.spill_stride = static_cast<int>(gridDim.x * blockDim.x);
  st.lds_depth = lds_depth;
  st.sp = 0;
  st.overflow = false;
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

__device__ __forceinline__ Surface resolve(const DevScene& S, const HitRef& h, f3 o, f3 d) {
  const InstRec in = load_inst(S.inst, h.inst);
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

// Gradient sky (scene/environment.h:34-46).
__device__ __forceinline__ f3 sky_eval(const DevSky& k, f3 d) {
  const f3 hz{k.horizon[0], k.horizon[1], k.horizon[2]};
  if (d.y < 0.f) {
    const float t = smin(-d.y, 1.f);
    return k.intensity * lerp3(hz, f3{k.ground[0], k.ground[1], k.ground[2]}, t);
  }
  const float t = powf(d.y, 1.0f / k.sharpness);
  return k.intensity * lerp3(hz, f3{k.zenith[0], k.zenith[1], k.zenith[2]}, t);
}

}  // namespace

// ---------------------------------------------------------------------------
// Batch bookkeeping shared by the wavefront kernels.
// ---------------------------------------------------------------------------
namespace {

// Per-block copy of the current batch's chunk table with prefix sums of
// 64-entry work units (traversal / shading) or of pixels (accumulation).
struct BatchView {
  int batch, nch;
  DevChunk ch[kMaxChunks];
  uint32_t cnt[kMaxChunks];
  uint32_t prefix[kMaxChunks + 1];
};

// Loads the batch and builds prefix[] from per-chunk counts (counts may be
// null: prefix over pixels). Block-wide; ends with a barrier.
__device__ __forceinline__ void load_batch(BatchView& v, const WaveBuffers& B, const uint32_t* counts,
                                           bool pixels) {
  if (threadIdx.x == 0) {
    v.batch = static_cast<int>(B.counters[kBatch]);
    v.nch = B.nchunks[v.batch];
  }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < v.nch) {
    v.ch[threadIdx.x] = B.chunks[v.batch * kMaxChunks + threadIdx.x];
    v.cnt[threadIdx.x] = counts ? counts[threadIdx.x] : 0u;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    v.prefix[0] = 0;
    for (int c = 0; c < v.nch; ++c) {
      acc += pixels ? static_cast<uint32_t>(v.ch[c].npix) : ((v.cnt[c] + 63u) >> 6);
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

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

}  // namespace

// ---------------------------------------------------------------------------
// raygen: per-(pixel, sample) seed and jittered pinhole ray
// (cpu_tiled.cpp:20-29, camera/camera.h:35-37); resets the batch's queue
// counters and work cursors.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_raygen(WaveBuffers B) {
  __shared__ BatchView v;
  if (threadIdx.x == 0) {
    v.batch = static_cast<int>(B.counters[kBatch]);
    v.nch = B.nchunks[v.batch];
  }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < v.nch) v.ch[threadIdx.x] = B.chunks[v.batch * kMaxChunks + threadIdx.x];
  __syncthreads();
  const int gid = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (gid < v.nch) {
    const DevChunk& ck = v.ch[gid];
    B.counters[kCnt + gid] = B.jobs[ck.job].maxb > 0 ? static_cast<uint32_t>(ck.cap) : 0u;
    B.counters[kCnt + kMaxChunks + gid] = 0u;
    B.counters[kShCnt + gid] = 0u;
  }
  if (gid < kNumCursors) B.counters[kCursor + gid] = 0u;
  if (v.nch == 0) return;
  const DevChunk& last = v.ch[v.nch - 1];
  if (gid >= last.base + last.cap) return;
  int lo = 0, hi = v.nch;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (v.ch[mid].base <= gid) lo = mid;
    else hi = mid;
  }
  const DevChunk& ck = v.ch[lo];
  const DevJob& J = B.jobs[ck.job];
  const int local = gid - ck.base;
  const int s = local / ck.npix;
  const int pixel = local - s * ck.npix;
  const int spp = ck.spp_first + s;
  const uint64_t seed = static_cast<uint64_t>(pixel) + static_cast<uint64_t>(spp) * 0xdeadbeefULL;
  Rng rng = rng_for(seed);
  const int x = pixel % J.W, y = pixel / J.W;
  const float fx = static_cast<float>(x) + rng.f();
  const float fy = static_cast<float>(y) + rng.f();
  const float u = fx / static_cast<float>(J.W), w = fy / static_cast<float>(J.H);
  const f3 eye{J.cam.eye[0], J.cam.eye[1], J.cam.eye[2]};
  const f3 llc{J.cam.llc[0], J.cam.llc[1], J.cam.llc[2]};
  const f3 hz{J.cam.horiz[0], J.cam.horiz[1], J.cam.horiz[2]};
  const f3 vt{J.cam.vert[0], J.cam.vert[1], J.cam.vert[2]};
  const f3 d = normalize(llc + u * hz + w * vt - eye);
  const PathQueue& q = B.q[0];
  q.ray0[gid] = make_float4(eye.x, eye.y, eye.z, d.x);
  q.ray1[gid] = make_float4(d.y, d.z, 0.f, __uint_as_float(1u));
  q.thr[gid] = make_float4(1.f, 1.f, 1.f, 0.f);
  q.st[gid] = make_uint4(static_cast<uint32_t>(rng.s), static_cast<uint32_t>(rng.s >> 32), static_cast<uint32_t>(gid),
                         0u);
  B.l_term[gid] = make_float4(0.f, 0.f, 0.f, 0.f);
  B.l_nee[2 * gid] = make_float4(0.f, 0.f, 0.f, 0.f);
  B.l_nee[2 * gid + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------------------
// Persistent traversal of one queue with per-lane refill. Waves take
// 64-entry units from a device cursor into a wave-local pool; a lane whose
// ray is finished takes the next pool entry at once, so a wave keeps all
// lanes busy until the queue runs dry instead of waiting for its slowest
// ray. A unit lies in one chunk; the rays in flight in a wave always share
// one job (and scene): a unit of another job waits until the wave drains.
// seg: queue entries per chunk sample slot (1 path queue, 2 shadow queue).
// Must be reached by every lane of the wave.
// ---------------------------------------------------------------------------
template <bool kAny, class Load, class Finish>
__device__ __forceinline__ void trace_queue(const WaveBuffers& B, const BatchView& v, uint32_t* cursor,
                                            uint32_t seg, TravStack& st, TravCount& c, uint32_t& rays,
                                            bool& overflow, Load load, Finish finish) {
  constexpr uint32_t kIdle = 0xFFFFFFFFu;
  const uint32_t total = v.prefix[v.nch];
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t my = kIdle;
  Walker<kAny> w;
  uint32_t pnext = 0, pend = 0;   // wave-uniform pool of offsets in chunk pci
  int pci = 0;
  int held = -1;                  // fetched unit of another job, waiting for the wave to drain
  bool exhausted = false;
  for (;;) {
    const unsigned long long act = __ballot(my != kIdle);
    if (act != ~0ull) {
      if (pnext >= pend) {
        int unit = -1;
        if (held >= 0) {
          if (act == 0ull) {
            unit = held;
            held = -1;
          }
        } else if (!exhausted) {
          uint32_t u = 0;
          if (lane == 0) u = atomicAdd(cursor, 1u);
          u = static_cast<uint32_t>(uniform(static_cast<int>(__shfl(u, 0))));
          if (u >= total) exhausted = true;
          else if (act != 0ull && v.ch[find_chunk(v, u)].job != v.ch[pci].job) held = static_cast<int>(u);
          else unit = static_cast<int>(u);
        }
        if (unit >= 0) {
          pci = uniform(find_chunk(v, static_cast<uint32_t>(unit)));
          pnext = (static_cast<uint32_t>(unit) - v.prefix[pci]) * 64u;
          pend = min(pnext + 64u, v.cnt[pci]);
        }
      }
      if (pnext < pend) {
        const unsigned long long idle = ~act;
        const uint32_t rank = static_cast<uint32_t>(__popcll(idle & below));
        const DevScene& S = B.jobs[uniform(v.ch[pci].job)].S;
        if (my == kIdle && pnext + rank < pend) {
          my = seg * static_cast<uint32_t>(v.ch[pci].base) + pnext + rank;
          f3 o, d;
          float tmin, tmax;
          load(my, o, d, tmin, tmax);
          w.begin(S, o, d, tmin, tmax, st);
        }
        pnext = min(pend, pnext + static_cast<uint32_t>(__popcll(idle)));
      }
    }
    if (__ballot(my != kIdle) == 0ull) break;
    if (my != kIdle) {
      const DevScene& S = B.jobs[uniform(v.ch[pci].job)].S;
      if (w.step(S, st, c)) {
        overflow |= st.overflow;
        finish(my, w);
        ++rays;
        my = kIdle;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// closest: persistent closest-hit traversal of queue qi.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_closest(WaveBuffers B, int qi, int cursor, int lds_depth) {
  extern __shared__ uint32_t lds[];
  __shared__ BatchView v;
  const uint32_t* counts = B.counters + kCnt + qi * kMaxChunks;
  load_batch(v, B, counts, false);
  if (blockIdx.x == 0 && static_cast<int>(threadIdx.x) < v.nch) {
    B.counters[kCnt + (1 - qi) * kMaxChunks + threadIdx.x] = 0u;
    B.counters[kShCnt + threadIdx.x] = 0u;
  }
  const PathQueue q = B.q[qi];
  TravStack st = make_stack(lds, B.spill, lds_depth);
  TravCount c{0, 0, 0, 0};
  uint32_t rays = 0;
  bool overflow = false;
  trace_queue<false>(
      B, v, &B.counters[kCursor + cursor], 1u, st, c, rays, overflow,
      [&](uint32_t i, f3& o, f3& d, float& tmin, float& tmax) {
        const float4 r0 = q.ray0[i], r1 = q.ray1[i];
        o = f3{r0.x, r0.y, r0.z};
        d = f3{r0.w, r1.x, r1.y};
        tmin = kBounceOffset;
        tmax = __builtin_huge_valf();
      },
      [&](uint32_t i, const Walker<false>& w) {
        B.hit[i] = make_float4(w.hit.t, w.hit.u, w.hit.v, __int_as_float(w.found ? w.hit.inst : -1));
        B.hit_slot[i] = w.hit.slot;
      });
  flush_counts(B.stats, c, rays, false);
  if (overflow) atomicAdd(&B.counters[kOverflow], 1u);
}

// ---------------------------------------------------------------------------
// shade: emission + MIS, area and sun NEE, Russian roulette, BSDF sampling
// (cpu_tiled.cpp:30-216) for every entry of queue qi at depth `depth`.
// Survivors and shadow rays are appended to the same chunk's segments.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_shade(WaveBuffers B, int qi, int depth) {
  __shared__ BatchView v;
  const uint32_t* counts = B.counters + kCnt + qi * kMaxChunks;
  load_batch(v, B, counts, false);
  const uint32_t total = v.prefix[v.nch];
  const PathQueue q = B.q[qi];
  const PathQueue qn = B.q[1 - qi];
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t unit = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); unit < total; unit += waves) {
    const int ci = uniform(find_chunk(v, unit));
    const DevChunk& ck = v.ch[ci];
    const DevJob& J = B.jobs[uniform(ck.job)];
    const DevScene& S = J.S;
    const uint32_t off = (unit - v.prefix[ci]) * 64u + static_cast<uint32_t>(lane_id());
    const uint32_t i = static_cast<uint32_t>(ck.base) + off;
    bool want_a = false, want_s = false, want_n = false;
    f3 ao{}, ad{}, ac{}, so{}, sd{}, sc{};
    float atmax = 0.f;
    f3 no{}, nd{}, nT{};
    float npdf = 0.f;
    uint32_t ndelta = 0u;
    uint32_t slot = 0u;
    Rng rng{0};
    if (off < v.cnt[ci]) {
      const float4 r0 = q.ray0[i], r1 = q.ray1[i], th = q.thr[i];
      const uint4 ps = q.st[i];
      const f3 o{r0.x, r0.y, r0.z}, d{r0.w, r1.x, r1.y};
      const float prev_pdf = r1.z;
      const bool prev_delta = (__float_as_uint(r1.w) & 1u) != 0u;
      f3 T{th.x, th.y, th.z};
      rng.s = static_cast<uint64_t>(ps.x) | (static_cast<uint64_t>(ps.y) << 32);
      slot = ps.z;
      const float4 hv = B.hit[i];
      HitRef h{hv.x, hv.y, hv.z, B.hit_slot[i], __float_as_int(hv.w)};
      if (h.inst < 0) {
        // miss: sky, no MIS, path ends (cpu_tiled.cpp:36-56)
        if (S.sky.intensity > 0.f) {
          const f3 L = T * sky_eval(S.sky, d);
          B.l_term[slot] = make_float4(L.x, L.y, L.z, 0.f);
        }
      } else {
        const Surface sf = resolve(S, h, o, d);
        int li = -1;
        if (sf.lmap >= 0) li = S.lmap[sf.lmap + static_cast<int>(sf.tri)];
        if (li >= 0) {
          // emitter hit with power-2 MIS against the previous BSDF sample;
          // the path ends (cpu_tiled.cpp:58-86)
          const float4 l0 = S.lights[4 * li], l1 = S.lights[4 * li + 1], l2 = S.lights[4 * li + 2],
                       l3 = S.lights[4 * li + 3];
          const f3 em{l1.w, l2.w, l3.w};
          float w = 1.0f;
          const float cl = smax(0.0f, dot(sf.gn, -d));
          if (cl > 1e-6f) {
            if (depth > 0 && !prev_delta) {
              const float pa = (1.0f / l0.w) * (1.0f / static_cast<float>(S.num_lights));
              const float d2 = h.t * h.t;
              const float psa = (pa * d2) / cl;
              w = mis_power2(prev_pdf, psa);
            }
            const f3 L = T * em * w;
            B.l_term[slot] = make_float4(L.x, L.y, L.z, 0.f);
          }
        } else {
          const DMat m = load_mat(S.mats, sf.mat);
          const bool delta = m.delta();
          // area-light NEE (cpu_tiled.cpp:88-139)
          if (S.num_lights > 0 && !delta) {
            const int nl = S.num_lights;
            const float up = rng.f();
            int k = static_cast<int>(up * static_cast<float>(nl));
            k = k < 0 ? 0 : (k > nl - 1 ? nl - 1 : k);
            const float pick = 1.f / static_cast<float>(nl);
            const float4 l0 = S.lights[4 * k], l1 = S.lights[4 * k + 1], l2 = S.lights[4 * k + 2],
                         l3 = S.lights[4 * k + 3];
            const float u1 = rng.f();
            const float u2 = rng.f();
            const float su1 = sqrtf(u1);
            const float b0 = 1.0f - su1;
            const float b1 = u2 * su1;
            const float b2 = 1.0f - b0 - b1;
            const f3 lp = xyz(l0) * b0 + xyz(l1) * b1 + xyz(l2) * b2;
            const f3 sdv = lp - sf.pos;
            const float d2 = len2(sdv);
            const float dist = sqrtf(d2);
            if (dist > 1e-6f) {
              const f3 wi = sdv / dist;
              const float cl = smax(0.0f, dot(xyz(l3), -wi));
              const Frame fr(sf.gn);
              const f3 wo_l = fr.local(-d);
              const f3 wi_l = fr.local(wi);
              if (cl > 1e-6f && wi_l.z > 1e-6f) {
                const f3 f = nee_f(wo_l, wi_l, m);
                const float bp = nee_pdf(wo_l, wi_l, m);
                const float lpdf = ((1.0f / l0.w) * pick) * (d2 / cl);
                const float w = mis_power2(lpdf, bp);
                const float cs = wi_l.z;
                if (w > 0.0f && lpdf > 1e-12f) {
                  want_a = true;
                  ao = offset_origin(sf.pos, sf.gn, wi, kShadowOff);
                  ad = wi;
                  atmax = dist - kShadowOff;
                  ac = T * f * cs * (f3{l1.w, l2.w, l3.w} / lpdf) * w;
                }
              }
            }
          }
          // sun NEE with the reference's 1e8 pdf sentinel (cpu_tiled.cpp:148-187)
          if (S.num_suns > 0 && !delta) {
            const int ns = S.num_suns;
            const float us = rng.f();
            int k = static_cast<int>(us * static_cast<float>(ns));
            k = k < 0 ? 0 : (k > ns - 1 ? ns - 1 : k);
            const float spick = 1.f / static_cast<float>(ns);
            if (spick > 0.f) {
              const f3 sdir = xyz(S.suns[2 * k]);
              const f3 sem = xyz(S.suns[2 * k + 1]);
              const Frame fr(sf.gn);
              const f3 wo_l = fr.local(-d);
              const f3 wi_l = fr.local(sdir);
              if (wi_l.z > 1e-6f) {
                const f3 f = nee_f(wo_l, wi_l, m);
                const float bp = nee_pdf(wo_l, wi_l, m);
                const float cs = wi_l.z;
                const float lpdf = 1e8f * spick;
                const float w = mis_power2(lpdf, bp);
                want_s = true;
                so = offset_origin(sf.pos, sf.gn, sdir, kShadowOff);
                sd = sdir;
                sc = T * f * cs * (sem / lpdf) * w;
              }
            }
          }
          // The last segment's continuation is never traced: its roulette
          // and BSDF draws cannot change the radiance, so they are skipped.
          if (depth < J.maxb - 1) {
            bool alive = true;
            if (depth >= J.minb) {   // Russian roulette (cpu_tiled.cpp:189-194)
              const float pm = smax(0.05f, smax(smax(T.x, T.y), T.z));
              if (rng.f() > pm) alive = false;
              else T = T * (1.f / pm);
            }
            if (alive) {
              const Frame fs(sf.n);
              const f3 wo_l = fs.local(-d);
              BSample bs;
              if (!sample_bsdf(wo_l, m, rng, bs)) {
                alive = false;
              } else if (bs.delta) {
                T = T * bs.f;
              } else {
                const float ci2 = fabsf(bs.wi.z);
                if (bs.pdf < 1e-8f || ci2 < 1e-8f) alive = false;
                else T = T * (bs.f * ci2 / bs.pdf);
              }
              if (alive) {
                const f3 wiw = fs.world(bs.wi);
                want_n = true;
                no = offset_origin(sf.pos, sf.gn, wiw, kBounceOffset);
                nd = wiw;
                nT = T;
                npdf = bs.pdf;
                ndelta = bs.delta ? 1u : 0u;
              }
            }
          }
        }
      }
    }
    // convergent section: compacted appends into this chunk's segments
    const uint32_t sbase = 2u * static_cast<uint32_t>(ck.base);
    const uint32_t ia = sbase + wave_append(&B.counters[kShCnt + ci], want_a);
    const uint32_t is = sbase + wave_append(&B.counters[kShCnt + ci], want_s);
    const uint32_t in = static_cast<uint32_t>(ck.base) +
                        wave_append(&B.counters[kCnt + (1 - qi) * kMaxChunks + ci], want_n);
    if (want_a) {
      B.sh0[ia] = make_float4(ao.x, ao.y, ao.z, ad.x);
      B.sh1[ia] = make_float4(ad.y, ad.z, atmax, __uint_as_float(slot * 2u));
      B.sh_c[ia] = make_float4(ac.x, ac.y, ac.z, 0.f);
    }
    if (want_s) {
      B.sh0[is] = make_float4(so.x, so.y, so.z, sd.x);
      B.sh1[is] = make_float4(sd.y, sd.z, __builtin_huge_valf(), __uint_as_float(slot * 2u + 1u));
      B.sh_c[is] = make_float4(sc.x, sc.y, sc.z, 0.f);
    }
    if (want_n) {
      qn.ray0[in] = make_float4(no.x, no.y, no.z, nd.x);
      qn.ray1[in] = make_float4(nd.y, nd.z, npdf, __uint_as_float(ndelta));
      qn.thr[in] = make_float4(nT.x, nT.y, nT.z, 0.f);
      qn.st[in] = make_uint4(static_cast<uint32_t>(rng.s), static_cast<uint32_t>(rng.s >> 32), slot, 0u);
    }
  }
}

// ---------------------------------------------------------------------------
// shadow: persistent any-hit traversal; unoccluded contributions are added
// to their path's per-slot NEE sum (one writer per (slot, light kind)).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_shadow(WaveBuffers B, int cursor, int lds_depth) {
  extern __shared__ uint32_t lds[];
  __shared__ BatchView v;
  const uint32_t* counts = B.counters + kShCnt;
  load_batch(v, B, counts, false);
  TravStack st = make_stack(lds, B.spill, lds_depth);
  TravCount c{0, 0, 0, 0};
  uint32_t rays = 0;
  bool overflow = false;
  trace_queue<true>(
      B, v, &B.counters[kCursor + cursor], 2u, st, c, rays, overflow,
      [&](uint32_t i, f3& o, f3& d, float& tmin, float& tmax) {
        const float4 r0 = B.sh0[i], r1 = B.sh1[i];
        o = f3{r0.x, r0.y, r0.z};
        d = f3{r0.w, r1.x, r1.y};
        tmin = kShadowOff;
        tmax = r1.z;
      },
      [&](uint32_t i, const Walker<true>& w) {
        if (!w.found) {
          const uint32_t k = __float_as_uint(B.sh1[i].w);
          const float4 cv = B.sh_c[i];
          float4 acc = B.l_nee[k];
          acc.x = acc.x + cv.x;
          acc.y = acc.y + cv.y;
          acc.z = acc.z + cv.z;
          B.l_nee[k] = acc;
        }
      });
  flush_counts(B.stats, c, rays, true);
  if (overflow) atomicAdd(&B.counters[kOverflow], 1u);
}

// ---------------------------------------------------------------------------
// accumulate: per (chunk, pixel), add each sample's path radiance to the
// job's sums in sample order (cpu_tiled.cpp:218). NaN/Inf are kept, as the
// reference keeps them, and counted.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_accumulate(WaveBuffers B) {
  __shared__ BatchView v;
  load_batch(v, B, nullptr, true);
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long bad = 0;
  unsigned long long paths = 0;
  if (gid < v.prefix[v.nch]) {
    const int ci = find_chunk(v, gid);
    const DevChunk& ck = v.ch[ci];
    const int p = static_cast<int>(gid - v.prefix[ci]);
    float4* accum = B.jobs[ck.job].accum;
    float4 a = accum[p];
    for (int s = 0; s < ck.nsamples; ++s) {
      const int slot = ck.base + s * ck.npix + p;
      const float4 t = B.l_term[slot], n0 = B.l_nee[2 * slot], n1 = B.l_nee[2 * slot + 1];
      const f3 L = f3{t.x, t.y, t.z} + f3{n0.x, n0.y, n0.z} + f3{n1.x, n1.y, n1.z};
      if (!finite3(L)) ++bad;
      a.x = a.x + L.x;
      a.y = a.y + L.y;
      a.z = a.z + L.z;
    }
    accum[p] = a;
    paths = static_cast<unsigned long long>(ck.nsamples);
  }
  const unsigned long long tot = wave_sum(bad), np = wave_sum(paths);
  if (lane_id() == 0) {
    if (tot) atomicAdd(&B.stats[kStNonFinite], tot);
    if (np) atomicAdd(&B.stats[kStPaths], np);
  }
}

// Moves the device to the next batch; the last node of a batch.
__global__ void k_advance(uint32_t* counters) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counters[kBatch] = counters[kBatch] + 1u;
}

__global__ void __launch_bounds__(256) k_average(const float4* accum, float* out, int npix, float inv_spp) {
  const int p = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= npix) return;
  const float4 a = accum[p];
  out[3 * p] = a.x * inv_spp;
  out[3 * p + 1] = a.y * inv_spp;
  out[3 * p + 2] = a.z * inv_spp;
}

// ---------------------------------------------------------------------------
// Ray-batch parity kernels (Scene::intersect / Scene::intersects).
// rays: 2 float4 per ray {o, tmin} {d, tmax}; hits: 16 words (xnh_hit).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_api_closest(DevScene S, const float4* rays, float* hits, int n,
                                                     uint32_t* spill, int lds_depth, unsigned long long* stats,
                                                     uint32_t* overflow_ctr) {
  extern __shared__ uint32_t lds[];
  TravStack st = make_stack(lds, spill, lds_depth);
  TravCount c{0, 0, 0, 0};
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  uint32_t rays_done = 0;
  if (i < n) {
    const float4 a = rays[2 * i], b = rays[2 * i + 1];
    const f3 o{a.x, a.y, a.z}, d{b.x, b.y, b.z};
    HitRef h{0.f, 0.f, 0.f, 0u, -1};
    const bool ok = walk_scene<false>(S, o, d, a.w, b.w, h, st, c);
    float* out = hits + 16 * static_cast<size_t>(i);
    if (!ok) {
      out[0] = __builtin_huge_valf();
      out[1] = 0.f;
      out[2] = 0.f;
      reinterpret_cast<int*>(out)[3] = -1;
      reinterpret_cast<int*>(out)[4] = -1;
      reinterpret_cast<int*>(out)[5] = -1;
      for (int k = 6; k < 15; ++k) out[k] = 0.f;
      reinterpret_cast<int*>(out)[15] = 0;
    } else {
      const Surface sf = resolve(S, h, o, d);
      out[0] = h.t;
      out[1] = h.u;
      out[2] = h.v;
      reinterpret_cast<int*>(out)[3] = sf.inst_id;
      reinterpret_cast<int*>(out)[4] = static_cast<int>(sf.tri);
      reinterpret_cast<int*>(out)[5] = sf.mat;
      out[6] = sf.pos.x; out[7] = sf.pos.y; out[8] = sf.pos.z;
      out[9] = sf.n.x; out[10] = sf.n.y; out[11] = sf.n.z;
      out[12] = sf.gn.x; out[13] = sf.gn.y; out[14] = sf.gn.z;
      reinterpret_cast<int*>(out)[15] = sf.front ? 1 : 0;
    }
    rays_done = 1;
    if (st.overflow) atomicAdd(overflow_ctr, 1u);
  }
  flush_counts(stats, c, rays_done, false);
}

__global__ void __launch_bounds__(256) k_api_any(DevScene S, const float4* rays, uint8_t* occ, int n, uint32_t* spill,
                                                 int lds_depth, unsigned long long* stats, uint32_t* overflow_ctr) {
  extern __shared__ uint32_t lds[];
  TravStack st = make_stack(lds, spill, lds_depth);
  TravCount c{0, 0, 0, 0};
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  uint32_t rays_done = 0;
  if (i < n) {
    const float4 a = rays[2 * i], b = rays[2 * i + 1];
    HitRef h{0.f, 0.f, 0.f, 0u, -1};
    occ[i] = walk_scene<true>(S, f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, a.w, b.w, h, st, c) ? 1 : 0;
    rays_done = 1;
    if (st.overflow) atomicAdd(overflow_ctr, 1u);
  }
  flush_counts(stats, c, rays_done, true);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int trace_grid(int block, int lds_depth) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
  int per_cu = 0;
  const size_t lds = static_cast<size_t>(block) * lds_depth * sizeof(uint32_t);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_closest), block, lds) !=
      hipSuccess)
    return -1;
  if (per_cu < 1) per_cu = 1;
  return prop.multiProcessorCount * per_cu;
}

int num_cus() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
  return prop.multiProcessorCount;
}

hipError_t launch_raygen(hipStream_t s, const WaveBuffers& B, int max_paths) {
  const int n = max_paths > kNumCursors ? max_paths : kNumCursors;
  hipLaunchKernelGGL(k_raygen, dim3((n + 255) / 256), dim3(256), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_closest(hipStream_t s, const WaveBuffers& B, int qi, int cursor, int grid, int block,
                          int lds_depth) {
  hipLaunchKernelGGL(k_closest, dim3(grid), dim3(block), static_cast<size_t>(block) * lds_depth * 4, s, B, qi, cursor,
                     lds_depth);
  return hipGetLastError();
}

hipError_t launch_shade(hipStream_t s, const WaveBuffers& B, int qi, int depth, int grid) {
  hipLaunchKernelGGL(k_shade, dim3(grid), dim3(256), 0, s, B, qi, depth);
  return hipGetLastError();
}

hipError_t launch_shadow(hipStream_t s, const WaveBuffers& B, int cursor, int grid, int block, int lds_depth) {
  hipLaunchKernelGGL(k_shadow, dim3(grid), dim3(block), static_cast<size_t>(block) * lds_depth * 4, s, B, cursor,
                     lds_depth);
  return hipGetLastError();
}

hipError_t launch_accumulate(hipStream_t s, const WaveBuffers& B, int max_pixels) {
  const int n = max_pixels > 1 ? max_pixels : 1;
  hipLaunchKernelGGL(k_accumulate, dim3((n + 255) / 256), dim3(256), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_advance(hipStream_t s, uint32_t* counters) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, counters);
  return hipGetLastError();
}

hipError_t launch_average(hipStream_t s, const float4* accum, float* out, int npix, float inv_spp) {
  hipLaunchKernelGGL(k_average, dim3((npix + 255) / 256), dim3(256), 0, s, accum, out, npix, inv_spp);
  return hipGetLastError();
}

hipError_t launch_api_closest(hipStream_t s, const DevScene& S, const float4* rays, float* hits, int n,
                              uint32_t* spill, int lds_depth, unsigned long long* stats, uint32_t* overflow_ctr) {
  const int block = 256;
  hipLaunchKernelGGL(k_api_closest, dim3((n + block - 1) / block), dim3(block),
                     static_cast<size_t>(block) * lds_depth * 4, s, S, rays, hits, n, spill, lds_depth, stats,
                     overflow_ctr);
  return hipGetLastError();
}

hipError_t launch_api_any(hipStream_t s, const DevScene& S, const float4* rays, uint8_t* occ, int n, uint32_t* spill,
                          int lds_depth, unsigned long long* stats, uint32_t* overflow_ctr) {
  const int block = 256;
  hipLaunchKernelGGL(k_api_any, dim3((n + block - 1) / block), dim3(block),
                     static_cast<size_t>(block) * lds_depth * 4, s, S, rays, occ, n, spill, lds_depth, stats,
                     overflow_ctr);
  return hipGetLastError();
}

}  // namespace xnh
