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
#include "dwave.h"
#include "guide.h"
#include "launch.h"

namespace xnh {

// ---------------------------------------------------------------------------
// raygen: starts a batch. Depth 0 of every chunk is its camera paths
// (segment 0 counts the chunk's paths; their order and the rays themselves
// come from camera_path, wherever a depth-0 entry is read); the batch's
// other queue and shadow counters, work cursors and record counters are
// reset and every slot's radiance terms cleared. Grid-stride over the
// batch's slots, so the grid is bounded whatever the queue capacity.
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
  const int tid = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (tid < v.nch * kQSeg) {
    // depth 0: the chunk's paths as its segment 0, in slot order
    const int ci = tid / kQSeg, x = tid % kQSeg;
    const DevChunk& ck = v.ch[ci];
    B.counters[ctr_path(0, ci, x)] = (x == 0 && B.jobs[ck.job].maxb > 0) ? static_cast<uint32_t>(ck.cap) : 0u;
    B.counters[ctr_path(1, ci, x)] = 0u;
    B.counters[ctr_shadow(ci, x)] = 0u;
  }
  if (tid < kNumCursors) B.counters[kCursor + tid] = 0u;
  if (tid < kQSeg) B.counters[ctr_rec(tid)] = 0u;
  if (v.nch == 0) return;
  // the chunks' slots are consecutive from 0
  const int end = v.ch[v.nch - 1].base + v.ch[v.nch - 1].cap;
  const int stride = static_cast<int>(gridDim.x * blockDim.x);
  for (int gid = tid; gid < end; gid += stride) {
    B.l_term[gid] = make_float4(0.f, 0.f, 0.f, 0.f);
    B.l_nee[2 * gid] = make_float4(0.f, 0.f, 0.f, 0.f);
    B.l_nee[2 * gid + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// ---------------------------------------------------------------------------
// Persistent traversal of one queue with per-lane refill. Waves take
// 64-entry units from a device cursor into a wave-local pool; a lane whose
// ray is finished takes the next pool entry at once, so a wave keeps all
// lanes busy until the queue runs dry instead of waiting for its slowest
// ray. A unit lies in one chunk; the rays in flight in a wave always share
// one job (and scene): a unit of another job waits until the wave drains.
// seg: queue entries per chunk sample slot (1 path queue, 2 shadow queue).
// kSort: the pool is all fetched units of one chunk (up to 16, 1024 rays)
// at once, ordered by the direction octant of their rays (bits 1-3 of
// ray1.w, set by the shading kernel), so that the rays a wave walks together
// visit the BVH's children in the same order; slist and sbase are the wave's
// LDS for the order and the units' first entries. Which rays share a wave
// changes no ray's result.
// Must be reached by every lane of the wave.
// ---------------------------------------------------------------------------
template <bool kAny, bool kCount, bool kPrims, bool kSort, class Load, class Finish>
__device__ __forceinline__ void trace_queue(const WaveBuffers& B, const BatchView& v, uint32_t* cursor,
                                            uint32_t seg, TravStack& st, TravCount& c, uint32_t& rays,
                                            bool& overflow, const float4* ray1, uint16_t* slist, uint32_t* sbase,
                                            Load load, Finish finish) {
  constexpr uint32_t kIdle = 0xFFFFFFFFu;
  const uint32_t total = v.prefix[v.nch];
  const int lane = lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t my = kIdle;
  Walker<kAny, kCount, kPrims> w;
  uint32_t pbase = 0, pnext = 0, pend = 0;   // wave-uniform pool: entries pbase + [pnext, pend) of chunk pci
  int pci = 0;
  // Units are taken from the cursor several at a time: one global atomic
  // on one address serves only ~85 M operations per second on the whole
  // device, so a fetch per 64-entry unit capped a launch at ~5 G rays/s.
  // A fetch takes up to 16 units, fewer as the queue runs out (what is left
  // per 4 waves of the grid, as this wave last saw it), so the launch
  // does not end with a few waves walking long private batches. Batches of
  // fewer than 64 units per wave (a 512x512 frame at 64 samples on the
  // whole device) walk short queues at every depth: there the fetches'
  // own cost outweighs the balance at the end of a launch, and a fetch
  // takes up to twice as much (what is left per 2 waves).
  const uint32_t gwaves = gridDim.x * (blockDim.x >> 6);
  const uint32_t batch_units =
      v.nch > 0 ? static_cast<uint32_t>(v.ch[v.nch - 1].base + v.ch[v.nch - 1].cap) >> 6 : 0u;
  const uint32_t share = batch_units < 64u * gwaves ? 2u : 4u;
  uint32_t seen = 0;              // the first unit of this wave's last fetch
  uint32_t ucur = 0, uend = 0;    // fetched units not yet taken (wave-uniform)
  bool exhausted = false;
  bool stuck = false;
  for (uint32_t guard = 0;; ++guard) {
    if (guard >= kWaveStepLimit) {
      stuck = true;
      break;
    }
    const unsigned long long act = __ballot(my != kIdle);
    if (act != ~0ull) {
      if (pnext >= pend) {
        if (ucur >= uend && !exhausted) {
          const uint32_t left = total > seen ? total - seen : 0u;
          const uint32_t per = max(1u, min(16u, left / (share * gwaves)));
          uint32_t u0 = 0;
          if (lane == 0) u0 = atomicAdd(cursor, per);
          u0 = static_cast<uint32_t>(uniform(static_cast<int>(__shfl(u0, 0))));
          if (u0 >= total) {
            exhausted = true;
          } else {
            ucur = u0;
            uend = min(u0 + per, total);
            seen = u0;
          }
        }
        if (ucur < uend) {
          // a unit of another job waits until the wave has drained (the
          // rays in flight in a wave share one scene)
          const int ci = uniform(find_chunk(v, ucur));
          if (act == 0ull || v.ch[ci].job == v.ch[pci].job) {
            pci = ci;
            if constexpr (kSort) {
              // the fetched units that lie in this chunk (1 to 16), keyed by
              // octant (4 bits per unit and lane; 8 = no entry)
              const uint32_t n = min(uend, v.prefix[ci + 1]) - ucur;
              uint32_t keys0 = 0u, keys1 = 0u;
              // every lane has taken its entry of the last pool before the
              // order is rewritten
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
              for (uint32_t j = 0; j < n; ++j) {
                uint32_t first, lim;
                const uint32_t b = unit_segment(v, ci, ucur + j - v.prefix[ci], seg, first, lim) + first;
                const uint32_t cnt = lim > first ? min(64u, lim - first) : 0u;
                if (lane == 0) sbase[j] = b;
                uint32_t key = 8u;
                if (static_cast<uint32_t>(lane) < cnt)
                  key = (reinterpret_cast<const uint32_t*>(ray1)[4u * (b + static_cast<uint32_t>(lane)) + 3u] >> 1) & 7u;
                if (j < 8u) keys0 |= key << (4u * j);
                else keys1 |= key << (4u * (j - 8u));
              }
              // counting sort: octant, then unit, then lane
              uint32_t pos = 0;
              for (uint32_t k = 0; k < 8u; ++k) {
                for (uint32_t j = 0; j < n; ++j) {
                  const uint32_t key = (j < 8u ? keys0 >> (4u * j) : keys1 >> (4u * (j - 8u))) & 15u;
                  const unsigned long long m = __ballot(key == k);
                  if (key == k) slist[pos + static_cast<uint32_t>(__popcll(m & below))] = static_cast<uint16_t>((j << 6) | static_cast<uint32_t>(lane));
                  pos += static_cast<uint32_t>(__popcll(m));
                }
              }
              // the order and the unit bases are read by other lanes
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
              pnext = 0;
              pend = pos;
              ucur += n;
            } else {
              uint32_t first, lim;
              pbase = unit_segment(v, ci, ucur - v.prefix[ci], seg, first, lim);
              pnext = first;
              pend = min(first + 64u, lim);
              ++ucur;
            }
          }
        }
      }
      if (pnext < pend) {
        const unsigned long long idle = ~act;
        const uint32_t rank = static_cast<uint32_t>(__popcll(idle & below));
        const DevScene& S = job_scene(B.jobs, uniform(v.ch[pci].job));
        if (my == kIdle && pnext + rank < pend) {
          if constexpr (kSort) {
            const uint32_t t = slist[pnext + rank];
            my = sbase[t >> 6] + (t & 63u);
          } else {
            my = pbase + pnext + rank;
          }
          f3 o, d;
          float tmin, tmax;
          load(my, pci, o, d, tmin, tmax);
          w.begin(S, o, d, tmin, tmax, st);
        }
        pnext = min(pend, pnext + static_cast<uint32_t>(__popcll(idle)));
      }
    }
    if (__ballot(my != kIdle) == 0ull) break;
    if (kCount && lane == 0) ++c.round_iters;
    if (kCount && my != kIdle) ++c.round_lanes;
    const DevScene& S = job_scene(B.jobs, uniform(v.ch[pci].job));
    if (wave_step(w, S, st, c, my != kIdle)) {
      overflow |= st.overflow;
      finish(my, w);
      ++rays;
      my = kIdle;
    }
  }
  if (__ballot(stuck) != 0ull && lane_id() == 0) atomicAdd(&B.counters[kWatchdog], 1u);
}

// ---------------------------------------------------------------------------
// closest: persistent closest-hit traversal of queue qi (kCount: with the
// traversal work counters; kCamera: depth 0, whose rays are the chunks'
// camera rays, rebuilt from the entry index).
// ---------------------------------------------------------------------------
// Per-wave LDS of the octant-ordered pool (bounce rays only).
template <bool kOn>
struct SortPool {
  uint16_t list[256 / 64][16 * 64];
  uint32_t base[256 / 64][16];
};
template <>
struct SortPool<false> {
  uint16_t list[1][1];
  uint32_t base[1][1];
};

template <bool kCount, bool kPrims, bool kCamera>
__global__ void __launch_bounds__(256) k_closest(WaveBuffers B, int qi, int cursor, int lds_depth) {
  __shared__ BatchView v;
  __shared__ SortPool<!kCamera> sp;
  const uint32_t* counts = B.counters + ctr_path(qi, 0, 0);
  load_batch(v, B, counts, false);
  if (blockIdx.x == 0)
    for (int t = static_cast<int>(threadIdx.x); t < v.nch * kQSeg; t += static_cast<int>(blockDim.x)) {
      B.counters[ctr_path(1 - qi, t / kQSeg, t % kQSeg)] = 0u;
      B.counters[ctr_shadow(t / kQSeg, t % kQSeg)] = 0u;
    }
  const PathQueue q = B.q[qi];
  TravStack st = make_stack(B.spill, lds_depth);
  TravCount c{0, 0, 0, 0};
  uint32_t rays = 0;
  bool overflow = false;
  const uint32_t wv = kCamera ? 0u : threadIdx.x >> 6;
  trace_queue<false, kCount, kPrims, !kCamera>(
      B, v, &B.counters[kCursor + cursor], 1u, st, c, rays, overflow, q.ray1, sp.list[wv], sp.base[wv],
      [&](uint32_t i, int ci, f3& o, f3& d, float& tmin, float& tmax) {
        if constexpr (kCamera) {
          const DevChunk& ck = v.ch[ci];
          const CameraPath p = camera_path(ck, job_rec(B.jobs, uniform(ck.job)), i);
          o = p.o;
          d = p.d;
        } else {
          const float4 r0 = q.ray0[i], r1 = q.ray1[i];
          o = f3{r0.x, r0.y, r0.z};
          d = f3{r0.w, r1.x, r1.y};
        }
        tmin = kBounceOffset;
        tmax = __builtin_huge_valf();
      },
      [&](uint32_t i, const Walker<false, kCount, kPrims>& w) {
        B.hit[i] = make_float4(w.found ? w.tbest : 0.f, w.hit.u, w.hit.v, __int_as_float(w.found ? w.hit.inst : -1));
        B.hit_slot[i] = w.hit.slot;
      });
  flush_counts_block(B.stats, c, rays, false, kCount);
  if (overflow) atomicAdd(&B.counters[kOverflow], 1u);
}

// ---------------------------------------------------------------------------
// shade: emission + MIS, area and sun NEE, Russian roulette, BSDF sampling
// (cpu_tiled.cpp:30-216) for every entry of queue qi at depth `depth`.
// Survivors and shadow rays are appended to the same chunk's segments.
//
// Each wave takes a window of kShadeWin consecutive 64-entry units and, when
// the window lies in one chunk, orders its entries by shading class (miss,
// Lambert, conductor, ...: the branch the entry takes below) before shading
// them 64 at a time, so that most of its passes run one branch with all
// lanes instead of every branch present with a few lanes each. Every
// entry's shading depends only on its own path, so the order changes
// neither its result nor where it is written (per-slot radiance terms;
// appended rays carry their slot).
// ---------------------------------------------------------------------------
// Units per shading window (kShadeWin). Eight sort better than four and
// amortise the window's class loads over more entries; at four waves per
// SIMD (128 VGPRs, a few bytes spilled) the plain kernel beat three waves
// without the spill (benchmark k_shade 225 -> 217 ms per two steps). The
// guided kernel keeps its three waves.

template <bool kGuided>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kGuided ? 3 : 4))) k_shade(WaveBuffers B, int qi, int depth) {
  __shared__ BatchView v;
  __shared__ uint16_t perm_all[256 / 64][kShadeWin * 64];   // per wave: window positions in class order
  __shared__ uint2 uinfo_all[256 / 64][kShadeWin];           // per wave: units' first queue entry, entry count
  const uint32_t* counts = B.counters + ctr_path(qi, 0, 0);
  load_batch(v, B, counts, false);
  const uint32_t total = v.prefix[v.nch];
  const PathQueue q = B.q[qi];
  const PathQueue qn = B.q[1 - qi];
  const uint32_t wv = threadIdx.x >> 6;
  uint16_t* perm = perm_all[wv];
  uint2* uinfo = uinfo_all[wv];
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t u0 = (blockIdx.x * (blockDim.x >> 6) + wv) * kShadeWin; u0 < total; u0 += waves * kShadeWin) {
    const uint32_t nu = min(kShadeWin, total - u0);
    const int ci0 = uniform(find_chunk(v, u0));
    const bool sorted = uniform(find_chunk(v, u0 + nu - 1)) == ci0;
    // every unit's first queue entry and entry count (wave-uniform; lane u
    // keeps unit u's in LDS for the passes)
    uint32_t upos[kShadeWin], ucnt[kShadeWin];
#pragma unroll
    for (uint32_t u = 0; u < kShadeWin; ++u) {
      upos[u] = 0u;
      ucnt[u] = 0u;
      if (u < nu) {
        const int cu = sorted ? ci0 : uniform(find_chunk(v, u0 + u));
        uint32_t first, lim;
        upos[u] = unit_segment(v, cu, u0 + u - v.prefix[cu], 1u, first, lim) + first;
        ucnt[u] = lim > first ? min(64u, lim - first) : 0u;
      }
      if (lane == u) uinfo[u] = make_uint2(upos[u], ucnt[u]);
    }
    if (sorted) {
      // counting sort of the window's entries by class: class, then unit,
      // then lane order
      const DevChunk& ck0 = v.ch[ci0];
      const DevScene& S0 = job_scene(B.jobs, uniform(ck0.job));
      // all hit loads of the window, then all class loads: two round trips
      // (an entry past the end reads the window's first, which exists)
      int inst[kShadeWin];
#pragma unroll
      for (uint32_t u = 0; u < kShadeWin; ++u)
        inst[u] = __float_as_int(B.hit[u < nu && lane < ucnt[u] ? upos[u] + lane : upos[0]].w);
      uint32_t key[kShadeWin];
#pragma unroll
      for (uint32_t u = 0; u < kShadeWin; ++u) {
        key[u] = kShadeInvalid;
        if (u < nu && lane < ucnt[u]) key[u] = inst[u] < 0 ? static_cast<uint32_t>(kShadeMiss) : S0.ikind[inst[u]];
      }
      uint32_t base = 0;
      for (uint32_t k = 0; k < kNumShadeClasses; ++k) {
#pragma unroll
        for (uint32_t u = 0; u < kShadeWin; ++u) {
          const unsigned long long m = __ballot(key[u] == k);
          if (key[u] == k) perm[base + static_cast<uint32_t>(__popcll(m & below))] = static_cast<uint16_t>(u * 64u + lane);
          base += static_cast<uint32_t>(__popcll(m));
        }
      }
    }
    // the window's unit table and positions are read by other lanes of this
    // wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t p = 0; p < nu; ++p) {
    uint32_t pu = p, pl = lane;
    if (sorted) {
      const uint32_t pos = perm[p * 64u + lane];
      pu = pos >> 6;
      pl = pos & 63u;
    }
    const int ci = sorted ? ci0 : uniform(find_chunk(v, u0 + p));
    const DevChunk& ck = v.ch[ci];
    const DevJob& J = job_rec(B.jobs, uniform(ck.job));
    const DevScene& S = J.S;
    const uint2 ui = uinfo[pu];
    const uint32_t i = ui.x + pl;   // queue entry
    bool want_a = false, want_s = false, want_n = false;
    f3 ao{}, ad{}, ac{}, so{}, sd{}, sc{};
    float atmax = 0.f;
    f3 no{}, nd{}, nT{};
    float npdf = 0.f;
    uint32_t ndelta = 0u;
    uint32_t slot = 0u;
    uint32_t nrec = kNoRecord;
    Rng rng{0};
    if (pl < ui.y) {
      f3 o, d, T;
      float prev_pdf;
      bool prev_delta;
      uint32_t prec = kNoRecord;   // the training record the previous vertex started
      if (depth == 0) {
        // a camera path: unit throughput, no previous vertex (the emitter
        // weight below is 1 at depth 0)
        const CameraPath cp = camera_path(ck, J, i);
        o = cp.o;
        d = cp.d;
        T = f3{1.f, 1.f, 1.f};
        prev_pdf = 0.f;
        prev_delta = true;
        rng = cp.rng;
        slot = cp.slot;
      } else {
        const float4 r0 = q.ray0[i], r1 = q.ray1[i], th = q.thr[i];
        const uint4 ps = q.st[i];
        o = f3{r0.x, r0.y, r0.z};
        d = f3{r0.w, r1.x, r1.y};
        prev_pdf = r1.z;
        prev_delta = (__float_as_uint(r1.w) & 1u) != 0u;
        T = f3{th.x, th.y, th.z};
        rng.s = static_cast<uint64_t>(ps.x) | (static_cast<uint64_t>(ps.y) << 32);
        slot = ps.z;
        prec = ps.w;
      }
      if (kGuided && prec < B.rec_cap) {
        // the previous vertex recorded a training sample: note the path's
        // radiance so far (its NEE included) so that what follows is known
        const float4 a = B.l_nee[2 * slot], b = B.l_nee[2 * slot + 1];
        B.rec[3 * B.rec_cap + prec] = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, __uint_as_float(slot));
      }
      const float4 hv = B.hit[i];
      HitRef h{hv.x, hv.y, hv.z, B.hit_slot[i], __float_as_int(hv.w)};
      if (h.inst < 0) {
        // miss: sky, no MIS, path ends (cpu_tiled.cpp:36-56)
        if (S.sky.intensity > 0.f) {
          const f3 L = T * sky_eval(S.sky, d);
          B.l_term[slot] = make_float4(L.x, L.y, L.z, 0.f);
        }
      } else {
        const Surface sf = resolve(S, h, o, d, kBounceOffset);
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
          // guided vertex: one-sample MIS of the stored guide mixture and
          // the BSDF sampler, with probabilities alpha and 1 - alpha. At
          // depth 0 a pixel's samples share its first sample's mixture (see
          // k_guide_infer); an entry without a valid mixture is not guided.
          const float alpha = kGuided ? *B.alpha_now : 0.f;
          // depth 0: segment 0 in slot order, so entry - qbase is the
          // sample-major slot offset and the pixel's first sample is at
          // qbase + pixel
          const uint32_t mi = (kGuided && depth == 0)
                                  ? static_cast<uint32_t>(ck.qbase) + (i - static_cast<uint32_t>(ck.qbase)) % static_cast<uint32_t>(ck.npix)
                                  : i;
          const bool guided = kGuided && depth < B.guide_depth && lambert_exact(m) &&
                              B.mix[8 * B.mix_stride + mi].x >= 0.f;
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
                float bp = nee_pdf(wo_l, wi_l, m);
                if (guided) bp = alpha * stored_pdf(B.mix, B.mix_stride, mi, wi) + (1.f - alpha) * bp;
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
                float bp = nee_pdf(wo_l, wi_l, m);
                if (guided) bp = alpha * stored_pdf(B.mix, B.mix_stride, mi, sdir) + (1.f - alpha) * bp;
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
            if (alive && guided) {
              const Frame fs(sf.n);
              const f3 wo_l = fs.local(-d);
              f3 wi_l{0.f, 0.f, 0.f}, wiw{0.f, 0.f, 0.f};
              if (rng.f() < alpha) {
                const float u0 = rng.f();
                const float u1 = rng.f();
                const float u2 = rng.f();
                wiw = stored_sample(B.mix, B.mix_stride, mi, u0, u1, u2);
                wi_l = fs.local(wiw);
              } else {
                BSample bs;
                if (sample_bsdf(wo_l, m, rng, bs)) {
                  wi_l = bs.wi;
                  wiw = fs.world(bs.wi);
                }
              }
              // exact Lambertian: f = base (1 - metallic) / pi on wo's side
              // of the shading normal (the side the BSDF sampler draws on)
              const bool same = (wo_l.z < 0.f) ? (wi_l.z < 0.f) : (wi_l.z > 0.f);
              const float ci2 = fabsf(wi_l.z);
              const float pc = alpha * stored_pdf(B.mix, B.mix_stride, mi, wiw) + (1.f - alpha) * (same ? ci2 * kInvPiD : 0.f);
              if (!same || ci2 < 1e-8f || !(pc >= 1e-8f)) {
                alive = false;
              } else {
                const bool rec = B.rec_every > 0u && record_path(slot, v.batch, B.rec_every);
                // one atomic per wave, on the record segment of this window
                // (windows go to the segments in turn, so no segment takes
                // more than one window's entries over an equal share)
                const uint32_t rs = B.rec_cap / kQSeg, rx = (u0 / kShadeWin) % kQSeg;
                uint32_t r = wave_append(&B.counters[ctr_rec(static_cast<int>(rx))], rec);
                if (rec) {
                  if (r < rs) {
                    r += rx * rs;
                    const uint32_t R = B.rec_cap;
                    const float nan = __builtin_nanf("");
                    B.rec[r] = make_float4(sf.pos.x, sf.pos.y, sf.pos.z, T.x);
                    B.rec[R + r] = make_float4(sf.gn.x, sf.gn.y, sf.gn.z, T.y);
                    B.rec[2 * R + r] = make_float4(wiw.x, wiw.y, wiw.z, T.z);
                    B.rec[3 * R + r] = make_float4(nan, nan, nan, __uint_as_float(slot));
                    nrec = r;
                  }
                }
                T = T * (m.base * ((1.f - m.metallic) * kInvPiD * ci2 / pc));
                want_n = true;
                no = offset_origin(sf.pos, sf.gn, wiw, kBounceOffset);
                nd = wiw;
                nT = T;
                npdf = pc;
                ndelta = 0u;
              }
            } else if (alive) {
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
    // the work-group's segment of the chunk's next path queue and shadow
    // queue
    const uint32_t xs = blockIdx.x % kQSeg;
    uint32_t ia, is, in;
    wave_append_pass(&B.counters[ctr_shadow(ci, static_cast<int>(xs))], want_a, want_s, ia, is,
                     &B.counters[ctr_path(1 - qi, ci, static_cast<int>(xs))], want_n, in);
    {
      // a segment holds what its group can append (kSegSlack); an append
      // past its end would be an internal error: counted, dropped
      const uint32_t qseg = static_cast<uint32_t>(ck.qseg);
      const bool drop = (want_a && ia >= 2u * qseg) || (want_s && is >= 2u * qseg) || (want_n && in >= qseg);
      if (__ballot(drop) != 0ull) {
        if (lane == 0) atomicAdd(&B.counters[kSegOverflow], 1u);
        if (ia >= 2u * qseg) want_a = false;
        if (is >= 2u * qseg) want_s = false;
        if (in >= qseg) want_n = false;
      }
      const uint32_t segpos = static_cast<uint32_t>(ck.qbase) + xs * qseg;
      ia += 2u * segpos;
      is += 2u * segpos;
      in += segpos;
    }
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
      // flags: bit 0 delta, bits 1-3 the direction octant (the traversal
      // groups rays by it)
      const uint32_t oct = (nd.x < 0.f ? 1u : 0u) | (nd.y < 0.f ? 2u : 0u) | (nd.z < 0.f ? 4u : 0u);
      qn.ray1[in] = make_float4(nd.y, nd.z, npdf, __uint_as_float(ndelta | (oct << 1)));
      qn.thr[in] = make_float4(nT.x, nT.y, nT.z, 0.f);
      qn.st[in] = make_uint4(static_cast<uint32_t>(rng.s), static_cast<uint32_t>(rng.s >> 32), slot, nrec);
    }
    }
    // every lane has read its positions before the next window writes them
    __builtin_amdgcn_wave_barrier();
  }
}

// ---------------------------------------------------------------------------
// shadow: persistent any-hit traversal; unoccluded contributions are added
// to their path's per-slot NEE sum (one writer per (slot, light kind)).
// ---------------------------------------------------------------------------
template <bool kCount, bool kPrims>
__global__ void __launch_bounds__(256) k_shadow(WaveBuffers B, int cursor, int lds_depth) {
  __shared__ BatchView v;
  const uint32_t* counts = B.counters + ctr_shadow(0, 0);
  load_batch(v, B, counts, false, 2u);
  TravStack st = make_stack(B.spill, lds_depth);
  TravCount c{0, 0, 0, 0};
  uint32_t rays = 0;
  bool overflow = false;
  trace_queue<true, kCount, kPrims, false>(
      B, v, &B.counters[kCursor + cursor], 2u, st, c, rays, overflow, nullptr, nullptr, nullptr,
      [&](uint32_t i, int, f3& o, f3& d, float& tmin, float& tmax) {
        const float4 r0 = B.sh0[i], r1 = B.sh1[i];
        o = f3{r0.x, r0.y, r0.z};
        d = f3{r0.w, r1.x, r1.y};
        tmin = kShadowOff;
        tmax = r1.z;
      },
      [&](uint32_t i, const Walker<true, kCount, kPrims>& w) {
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
  flush_counts_block(B.stats, c, rays, true, kCount);
  if (overflow) atomicAdd(&B.counters[kOverflow], 1u);
}

// ---------------------------------------------------------------------------
// accumulate: per (chunk, pixel), add each sample's path radiance to the
// job's sums in sample order (cpu_tiled.cpp:218). NaN/Inf are kept, as the
// reference keeps them, and counted.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_accumulate(WaveBuffers B) {
  __shared__ BatchView v;
  __shared__ unsigned long long red[2][256 / 64];
  load_batch(v, B, nullptr, true);
  unsigned long long bad = 0;
  unsigned long long paths = 0;
  // grid-stride over the batch's pixels (the grid is sized for the largest
  // batch, most of whose entries have no pixel)
  const uint32_t total = v.prefix[v.nch];
  for (uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x; gid < total; gid += gridDim.x * blockDim.x) {
    const int ci = find_chunk(v, gid);
    const DevChunk& ck = v.ch[ci];
    const int p = static_cast<int>(gid - v.prefix[ci]);   // tiled pixel index
    const DevJob& J = B.jobs[ck.job];
    float4* accum = J.accum + tiled_pixel(p, J.W, J.H);
    float4 a = *accum;
    for (int s = 0; s < ck.nsamples; ++s) {
      const int slot = ck.base + s * ck.npix + p;
      const float4 t = B.l_term[slot], n0 = B.l_nee[2 * slot], n1 = B.l_nee[2 * slot + 1];
      const f3 L = f3{t.x, t.y, t.z} + f3{n0.x, n0.y, n0.z} + f3{n1.x, n1.y, n1.z};
      if (!finite3(L)) ++bad;
      a.x = a.x + L.x;
      a.y = a.y + L.y;
      a.z = a.z + L.z;
    }
    *accum = a;
    paths += static_cast<unsigned long long>(ck.nsamples);
  }
  // one atomic per block and counter
  const unsigned long long tb = wave_sum(bad), tp = wave_sum(paths);
  const int w = static_cast<int>(threadIdx.x >> 6);
  if (lane_id() == 0) {
    red[0][w] = tb;
    red[1][w] = tp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sb = 0, sp = 0;
    for (int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) {
      sb += red[0][k];
      sp += red[1][k];
    }
    if (sb) atomicAdd(&B.stats[kStNonFinite], sb);
    if (sp) atomicAdd(&B.stats[kStPaths], sp);
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
// rays: 2 float4 per ray in the xnh_ray layout {o.xyz, d.x} {d.yz, tmin,
// tmax}; hits: 16 words (xnh_hit).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_api_closest(DevScene S, const float4* rays, float* hits, int n,
                                                     uint32_t* spill, int lds_depth, unsigned long long* stats,
                                                     uint32_t* overflow_ctr) {
  TravStack st = make_stack(spill, lds_depth);
  TravCount c{0, 0, 0, 0};
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const bool valid = i < n;
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = make_float4(0.f, 1.f, 0.f, 0.f);
  if (valid) {
    a = rays[2 * i];
    b = rays[2 * i + 1];
  }
  const f3 o{a.x, a.y, a.z}, d{a.w, b.x, b.y};
  const float tmin = b.z, tmax = b.w;
  HitRef h{0.f, 0.f, 0.f, 0u, -1};
  const bool ok = walk_scene<false>(S, o, d, tmin, tmax, valid, h, st, c);
  if (valid) {
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
      float u = h.u, v = h.v;
      Surface sf;
      if (h.slot & kPrimHit) sf = resolve_prim(S, h, load_inst(S.inst, h.inst), o, d, tmin, u, v);
      else sf = resolve(S, h, o, d, tmin);
      out[0] = h.t;
      out[1] = u;
      out[2] = v;
      reinterpret_cast<int*>(out)[3] = sf.inst_id;
      reinterpret_cast<int*>(out)[4] = static_cast<int>(sf.tri);
      reinterpret_cast<int*>(out)[5] = sf.mat;
      out[6] = sf.pos.x; out[7] = sf.pos.y; out[8] = sf.pos.z;
      out[9] = sf.n.x; out[10] = sf.n.y; out[11] = sf.n.z;
      out[12] = sf.gn.x; out[13] = sf.gn.y; out[14] = sf.gn.z;
      reinterpret_cast<int*>(out)[15] = sf.front ? 1 : 0;
    }
    if (st.overflow) atomicAdd(overflow_ctr, 1u);
  }
  flush_counts(stats, c, valid ? 1u : 0u, false);
}

__global__ void __launch_bounds__(256) k_api_any(DevScene S, const float4* rays, uint8_t* occ, int n, uint32_t* spill,
                                                 int lds_depth, unsigned long long* stats, uint32_t* overflow_ctr) {
  TravStack st = make_stack(spill, lds_depth);
  TravCount c{0, 0, 0, 0};
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  const bool valid = i < n;
  float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = make_float4(0.f, 1.f, 0.f, 0.f);
  if (valid) {
    a = rays[2 * i];
    b = rays[2 * i + 1];
  }
  HitRef h{0.f, 0.f, 0.f, 0u, -1};
  const bool hit = walk_scene<true>(S, f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, b.z, b.w, valid, h, st, c);
  if (valid) {
    occ[i] = hit ? 1 : 0;
    if (st.overflow) atomicAdd(overflow_ctr, 1u);
  }
  flush_counts(stats, c, valid ? 1u : 0u, true);
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
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_closest<false, false, false>), block, lds) !=
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
  // at least kNumCursors threads (the counter resets), at most 4096
  // work-groups (grid-stride beyond)
  const int n = max_paths > kNumCursors ? max_paths : kNumCursors;
  const int blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
  hipLaunchKernelGGL(k_raygen, dim3(blocks), dim3(256), 0, s, B);
  return hipGetLastError();
}

template <bool kCount, bool kPrims>
void launch_closest_kind(hipStream_t s, const WaveBuffers& B, int qi, int cursor, int grid, int block, int lds_depth,
                         size_t lds, bool camera) {
  if (camera)
    hipLaunchKernelGGL((k_closest<kCount, kPrims, true>), dim3(grid), dim3(block), lds, s, B, qi, cursor, lds_depth);
  else
    hipLaunchKernelGGL((k_closest<kCount, kPrims, false>), dim3(grid), dim3(block), lds, s, B, qi, cursor, lds_depth);
}

hipError_t launch_closest(hipStream_t s, const WaveBuffers& B, int qi, int cursor, int grid, int block,
                          int lds_depth, bool count, bool prims, bool camera) {
  const size_t lds = static_cast<size_t>(block) * lds_depth * 4;
  if (prims) {
    if (count) launch_closest_kind<true, true>(s, B, qi, cursor, grid, block, lds_depth, lds, camera);
    else launch_closest_kind<false, true>(s, B, qi, cursor, grid, block, lds_depth, lds, camera);
  } else {
    if (count) launch_closest_kind<true, false>(s, B, qi, cursor, grid, block, lds_depth, lds, camera);
    else launch_closest_kind<false, false>(s, B, qi, cursor, grid, block, lds_depth, lds, camera);
  }
  return hipGetLastError();
}

// Past the guided depths the guided kernel takes the plain kernel's paths
// exactly: no entry is guided there, and the training records a guided vertex
// starts are completed by the next depth's shading. Those depths run the
// plain kernel, which keeps more waves per SIMD.
hipError_t launch_shade(hipStream_t s, const WaveBuffers& B, int qi, int depth, int grid) {
  if (B.guide_alpha > 0.f && depth <= B.guide_depth) hipLaunchKernelGGL(k_shade<true>, dim3(grid), dim3(256), 0, s, B, qi, depth);
  else hipLaunchKernelGGL(k_shade<false>, dim3(grid), dim3(256), 0, s, B, qi, depth);
  return hipGetLastError();
}

hipError_t launch_shadow(hipStream_t s, const WaveBuffers& B, int cursor, int grid, int block, int lds_depth,
                         bool count, bool prims) {
  const size_t lds = static_cast<size_t>(block) * lds_depth * 4;
  if (prims) {
    if (count) hipLaunchKernelGGL((k_shadow<true, true>), dim3(grid), dim3(block), lds, s, B, cursor, lds_depth);
    else hipLaunchKernelGGL((k_shadow<false, true>), dim3(grid), dim3(block), lds, s, B, cursor, lds_depth);
  } else {
    if (count) hipLaunchKernelGGL((k_shadow<true, false>), dim3(grid), dim3(block), lds, s, B, cursor, lds_depth);
    else hipLaunchKernelGGL((k_shadow<false, false>), dim3(grid), dim3(block), lds, s, B, cursor, lds_depth);
  }
  return hipGetLastError();
}

hipError_t launch_accumulate(hipStream_t s, const WaveBuffers& B, int max_pixels) {
  const int n = max_pixels > 1 ? max_pixels : 1;
  const int blocks = (n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048;
  hipLaunchKernelGGL(k_accumulate, dim3(blocks), dim3(256), 0, s, B);
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
