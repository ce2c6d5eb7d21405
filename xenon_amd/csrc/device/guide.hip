// guide.hip — training and evaluation kernels of the path-guiding network
// (see guide.h for the model): batched evaluation, the fused forward +
// KL-loss + backward pass with FP16 MFMA and FP32 accumulation, hash-grid
// gradient scatter, and Adam.
//
// Loss per training sample (Mueller et al. 2019): l = -lambda * log q(w),
// lambda = the sample's target weight (scaled by the caller so the batch
// mean is 1). Output gradients are derived in closed form:
//   softmax logit a_j:  -lambda (r_j - pi_j)
//   log kappa_j:        -lambda r_j kappa_j (1/k - 2/(e^{2k} - 1) + mu.w - 1)
//   raw mean m_j:       -lambda r_j kappa_j (w - (mu.w) mu) / |m_j|
// with responsibilities r_j = pi_j v_j(w) / q(w). Each per-sample output
// gradient is clipped to +-kGradClip so that it and the back-propagated
// gradients stay inside FP16 range (outlier samples with large lambda);
// back-propagated FP16 values saturate at +-kHalfMax. Gradients are summed
// over the batch (FP32 atomics); the caller divides by the batch size in Adam.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dbsdf.h"
#include "dwave.h"
#include "guide.h"
#include "launch.h"

namespace xnh {

namespace {

constexpr float kGradClip = 1024.f;
constexpr float kHalfMax = 65504.f;

__device__ __forceinline__ int glane() { return static_cast<int>(threadIdx.x & 63u); }

__device__ __forceinline__ unsigned long long g_wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ float g_wave_sumf(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Block-wide sum (blockDim.x a multiple of 64, at most 1024); every thread
// gets it. red: blockDim.x / 64 floats of LDS.
__device__ __forceinline__ float block_sumf(float v, float* red) {
  v = g_wave_sumf(v);
  const int w = static_cast<int>(threadIdx.x >> 6), nw = static_cast<int>(blockDim.x >> 6);
  if (glane() == 0) red[w] = v;
  __syncthreads();
  float t = 0.f;
  for (int k = 0; k < nw; ++k) t += red[k];
  __syncthreads();
  return t;
}

// Block-wide exclusive prefix sum of x; *total gets the block sum. red:
// blockDim.x / 64 + 1 words of LDS.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* red, uint32_t* total) {
  const int lane = glane(), w = static_cast<int>(threadIdx.x >> 6), nw = static_cast<int>(blockDim.x >> 6);
  uint32_t inc = x;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off);
    if (lane >= off) inc += y;
  }
  if (lane == 63) red[w] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int k = 0; k < nw; ++k) {
    if (k < w) before += red[k];
    all += red[k];
  }
  __syncthreads();
  *total = all;
  return before + inc - x;
}

// Output gradient of one sample; raw: 48 outputs; g: 48 FP32 gradients.
__device__ __forceinline__ float loss_grad(const float* raw, f3 w, float lam, float* g) {
  Mixture m;
  decode_mixture(raw, m);
  float v[kGLobes], q = 0.f;
  for (int j = 0; j < kGLobes; ++j) {
    v[j] = vmf_norm(m.kappa[j]) * expf(m.kappa[j] * (dot(m.mu[j], w) - 1.f));
    q += m.pi[j] * v[j];
  }
  q = fmaxf(q, 1e-20f);
  for (int k = 0; k < kGOut; ++k) g[k] = 0.f;
  for (int j = 0; j < kGLobes; ++j) {
    const float r = m.pi[j] * v[j] / q;
    const float k = m.kappa[j];
    const float dw = dot(m.mu[j], w);
    const float dlogv = 1.f / k - 2.f / expm1f(2.f * k) + (dw - 1.f);
    const float lk = raw[5 * j + 3];
    const float inside = (lk > kKappaLogMin && lk < kKappaLogMax) ? 1.f : 0.f;
    const float nm = sqrtf(raw[5 * j] * raw[5 * j] + raw[5 * j + 1] * raw[5 * j + 1] + raw[5 * j + 2] * raw[5 * j + 2] +
                           1e-12f);
    const f3 tang = w - dw * m.mu[j];
    const float c = -lam * r * k / nm;
    g[5 * j] = c * tang.x;
    g[5 * j + 1] = c * tang.y;
    g[5 * j + 2] = c * tang.z;
    g[5 * j + 3] = -lam * r * dlogv * k * inside;
    g[5 * j + 4] = -lam * (r - m.pi[j]);
  }
  for (int k = 0; k < kGLobes * 5; ++k) g[k] = fminf(fmaxf(g[k], -kGradClip), kGradClip);
  return -lam * logf(q);
}

// wave: dW (K x N, global FP32) = A^T (K x 64) * G (64 x N); A: 64 x K rows.
// dW is this work-group's own slice of the partial-gradient buffer (summed
// over work-groups by k_wgrad_reduce), so plain stores suffice.
__device__ __forceinline__ void weight_grad(const _Float16* A, int K, const _Float16* G, int N, float* dW) {
  const int lane = glane();
  for (int rt = 0; rt < K / 16; ++rt)
    for (int ct = 0; ct < N / 16; ++ct) {
      const f4v acc = mma_tile<true, false>(A, K, 16 * rt, G, N, 16 * ct, 64);
      for (int j = 0; j < 4; ++j) {
        const int r = 16 * rt + 4 * (lane >> 4) + j, col = 16 * ct + (lane & 15);
        dW[r * N + col] = acc[j];
      }
    }
}

// wave: Gout (64 x K) = (G (64 x N) * W^T) masked by (H > 0) when H != null;
// W: K x N row-major global. FP16 result, or FP32 into Gout32.
__device__ __forceinline__ void input_grad(const _Float16* G, int N, const _Float16* W, int K, const _Float16* H,
                                           _Float16* Gout, float* Gout32) {
  const int lane = glane();
  for (int rt = 0; rt < 4; ++rt)
    for (int ct = 0; ct < K / 16; ++ct) {
      const f4v acc = mma_tile<false, true>(G, N, 16 * rt, W, N, 16 * ct, N);
      for (int j = 0; j < 4; ++j) {
        const int r = 16 * rt + 4 * (lane >> 4) + j, col = 16 * ct + (lane & 15);
        float v = acc[j];
        if (H && !(static_cast<float>(H[r * K + col]) > 0.f)) v = 0.f;
        if (Gout) Gout[r * K + col] = static_cast<_Float16>(fminf(fmaxf(v, -kHalfMax), kHalfMax));
        else Gout32[r * K + col] = v;
      }
    }
  __syncthreads();
}

struct TrainLds {
  _Float16 X[64 * kGIn];
  _Float16 H1[64 * kGHid], H2[64 * kGHid], H3[64 * kGHid];
  _Float16 Ga[64 * kGHid], Gb[64 * kGHid];
  float F[64 * kGOut];
};

}  // namespace

// Evaluates the network for n points: raw outputs (n x 48, FP32).
__global__ void __launch_bounds__(64) k_guide_eval(GuideNet net, int n, const float* pos, const float* nrm,
                                                   float* raw) {
  __shared__ _Float16 lds[64 * kGIn];
  __shared__ __align__(16) h4v wl[kWTiles];
  stage_weights(net.wt, wl);
  const int lane = glane();
  const int i = static_cast<int>(blockIdx.x) * 64 + lane;
  const bool ok = i < n;
  const f3 p = ok ? f3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]} : f3{net.lo[0], net.lo[1], net.lo[2]};
  const f3 nn = ok ? f3{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]} : f3{0.f, 0.f, 1.f};
  guide_forward_wave(net, p, nn, lds, wl);
  const _Float16* out = lds + lane * kGIn;
  if (ok)
    for (int k = 0; k < kGOut; ++k) raw[static_cast<size_t>(i) * kGOut + k] = static_cast<float>(out[k]);
}

// Mixture queries for n points: with u non-null, draws one direction per
// point from its mixture with (u0, u1, u2) = u[3i..3i+2] and writes it to
// dir; then writes the mixture pdf at dir[i] (drawn or given) to pdf.
__global__ void __launch_bounds__(64) k_guide_query(GuideNet net, int n, const float* pos, const float* nrm,
                                                    const float* u, float* dir, float* pdf) {
  __shared__ _Float16 lds[64 * kGIn];
  __shared__ __align__(16) h4v wl[kWTiles];
  stage_weights(net.wt, wl);
  const int lane = glane();
  const int i = static_cast<int>(blockIdx.x) * 64 + lane;
  const bool ok = i < n;
  const f3 p = ok ? f3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]} : f3{net.lo[0], net.lo[1], net.lo[2]};
  const f3 nn = ok ? f3{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]} : f3{0.f, 0.f, 1.f};
  guide_forward_wave(net, p, nn, lds, wl);
  if (!ok) return;
  float raw[kGOut];
  for (int k = 0; k < kGLobes * 5; ++k) raw[k] = static_cast<float>(lds[lane * kGIn + k]);
  Mixture m;
  decode_mixture(raw, m);
  f3 w;
  if (u) {
    w = mixture_sample(m, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    dir[3 * i] = w.x;
    dir[3 * i + 1] = w.y;
    dir[3 * i + 2] = w.z;
  } else {
    w = f3{dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
  }
  pdf[i] = mixture_pdf(m, w);
}

// ---------------------------------------------------------------------------
// Guided rendering
// ---------------------------------------------------------------------------

// Evaluates the guide at every closest hit of queue qi that lands on a
// guided surface (an exact Lambertian that is not an emitter; the network's
// normal input is the geometric normal) and stores the mixture for k_shade.
// One wave per block, grid-stride over the batch's 64-entry units. Launched
// for the bounces that are guided (below the configured guide depth).
// Guide mixtures for the hits of queue qi (after its closest-hit pass), in
// two kernels: k_guide_encode gathers the hash-grid encoding of every entry
// whose hit is a non-emitting exact-Lambertian surface (memory-latency
// bound, so it runs lean and at high occupancy) into feat, and k_guide_mlp
// runs the MLP on those rows from LDS-resident weights and stores the
// mixture (MFMA bound). Every entry gets a mark in the first lobe weight:
// negative = no mixture, which the shading kernel checks. At depth 0 the
// queue is in camera-ray order (entry = sample * npix + pixel within a
// chunk) and only each pixel's first sample of the batch is evaluated: the
// pixel's other samples use that mixture as their guide too (a proposal that
// does not depend on their own random numbers, so the estimate stays
// unbiased). feat: 6 uint4 per entry, feat[k * C + i].
constexpr int kInferWaves = 4;   // waves per work-group of both kernels

// The entries of one 64-entry unit that need a mixture: false for units
// past the last evaluated entry (returns false for the whole unit).
struct InferUnit {
  uint32_t i, off, lim;
  const DevChunk* ck;
};
__device__ __forceinline__ bool infer_unit(const BatchView& v, uint32_t unit, int depth, InferUnit& u) {
  const int ci = uniform(find_chunk(v, unit));
  u.ck = &v.ch[ci];
  uint32_t first, lim;
  const uint32_t sbase = unit_segment(v, ci, unit - v.prefix[ci], 1u, first, lim);
  // depth 0: segment 0 holds the chunk's paths in slot order, the pixels'
  // first samples first
  u.lim = depth == 0 ? min(lim, static_cast<uint32_t>(u.ck->npix)) : lim;
  if (first >= u.lim) return false;   // uniform: a later sample of every pixel in the unit
  u.off = first + static_cast<uint32_t>(glane());
  u.i = sbase + u.off;
  return true;
}

__global__ void __launch_bounds__(64 * kInferWaves) k_guide_encode(WaveBuffers B, int qi, int depth, GuideNet net,
                                                                   uint4* feat) {
  __shared__ BatchView v;
  const uint32_t* counts = B.counters + ctr_path(qi, 0, 0);
  load_batch(v, B, counts, false);
  const uint32_t total = v.prefix[v.nch];
  const PathQueue q = B.q[qi];
  const uint32_t C = B.mix_stride;
  const uint32_t wid = threadIdx.x >> 6;
  for (uint32_t unit = blockIdx.x * kInferWaves + wid; unit < total; unit += gridDim.x * kInferWaves) {
    InferUnit u;
    if (!infer_unit(v, unit, depth, u)) continue;
    if (u.off >= u.lim) continue;
    const DevScene& S = job_scene(B.jobs, uniform(u.ck->job));
    bool want = false;
    f3 p{0.f, 0.f, 0.f}, n{0.f, 0.f, 1.f};
    const float4 hv = B.hit[u.i];
    const int inst = __float_as_int(hv.w);
    if (inst >= 0) {
      f3 o, d;
      if (depth == 0) {   // camera rays are not stored
        const CameraPath cp = camera_path(*u.ck, job_rec(B.jobs, uniform(u.ck->job)), u.i);
        o = cp.o;
        d = cp.d;
      } else {
        const float4 r0 = q.ray0[u.i], r1 = q.ray1[u.i];
        o = f3{r0.x, r0.y, r0.z};
        d = f3{r0.w, r1.x, r1.y};
      }
      const HitRef h{hv.x, hv.y, hv.z, B.hit_slot[u.i], inst};
      const GeoSurface sf = resolve_geo(S, h, o, d, kBounceOffset);
      const bool emitter = sf.lmap >= 0 && S.lmap[sf.lmap + static_cast<int>(sf.tri)] >= 0;
      if (!emitter && lambert_exact(load_mat(S.mats, sf.mat))) {
        want = true;
        p = sf.pos;
        n = sf.gn;
      }
    }
    B.mix[8 * C + u.i] = make_float4(want ? 1.f : -1.f, 0.f, 0.f, 0.f);
    if (!want) continue;
    uint32_t o[24];
    guide_encode_packed(net, p, n, o);
#pragma unroll
    for (int k = 0; k < 6; ++k) feat[k * C + u.i] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
  }
}

__global__ void __launch_bounds__(64 * kInferWaves) k_guide_mlp(WaveBuffers B, int qi, int depth, GuideNet net,
                                                                const uint4* feat) {
  __shared__ BatchView v;
  __shared__ _Float16 act[kInferWaves][64 * kGIn];
  __shared__ __align__(16) h4v wl[kWTiles];
  const uint32_t* counts = B.counters + ctr_path(qi, 0, 0);
  load_batch(v, B, counts, false);
  stage_weights(net.wt, wl);
  const uint32_t total = v.prefix[v.nch];
  const uint32_t C = B.mix_stride;
  const uint32_t wid = threadIdx.x >> 6;
  _Float16* a = act[wid];
  uint4* row = reinterpret_cast<uint4*>(a + glane() * kGIn);
  // waves work independently from here on (no work-group barrier below)
  for (uint32_t unit = blockIdx.x * kInferWaves + wid; unit < total; unit += gridDim.x * kInferWaves) {
    InferUnit u;
    if (!infer_unit(v, unit, depth, u)) continue;
    const bool want = u.off < u.lim && B.mix[8 * C + u.i].x > 0.f;
    if (__ballot(want) == 0ull) continue;
#pragma unroll
    for (int k = 0; k < 6; ++k) row[k] = want ? feat[k * C + u.i] : make_uint4(0u, 0u, 0u, 0u);
    guide_mlp_wave(a, wl);
    if (want) decode_store_mixture(a + glane() * kGIn, B.mix, C, u.i);
  }
}

// Lambda of record r (0 when its path never wrote the prefix).
// lambda = luminance((L - prefix) / T): the radiance the path gathered after
// the vertex over the throughput that reached it, i.e. f cos / p times the
// incident radiance along the sampled direction (the weight of the KL
// objective of Mueller et al. 2019).
__device__ __forceinline__ float record_lambda(const WaveBuffers& B, uint32_t r) {
  const uint32_t R = B.rec_cap;
  const float4 a = B.rec[r], b = B.rec[R + r], c = B.rec[2 * R + r], e = B.rec[3 * R + r];
  const uint32_t slot = __float_as_uint(e.w);
  const float4 t = B.l_term[slot], n0 = B.l_nee[2 * slot], n1 = B.l_nee[2 * slot + 1];
  const float ar = t.x + n0.x + n1.x - e.x, ag = t.y + n0.y + n1.y - e.y, ab = t.z + n0.z + n1.z - e.z;
  const float lr = a.w > 0.f ? ar / a.w : 0.f, lg = b.w > 0.f ? ag / b.w : 0.f, lb = c.w > 0.f ? ab / c.w : 0.f;
  const float l = 0.2126f * lr + 0.7152f * lg + 0.0722f * lb;
  return (l > 0.f && l < 3.0e38f) ? l : 0.f;
}

// Whether record r (lambda l) is drawn, and its weight: probability
// p = min(1, l / tau), weight l / p = max(l, tau).
__device__ __forceinline__ bool record_drawn(uint32_t r, float l, float tau, uint32_t seed, float& wgt) {
  if (!(l > 0.f)) return false;
  const float p = (tau > 0.f && l < tau) ? l / tau : 1.f;
  uint32_t h = r * 0x9E3779B1u ^ (seed * 0x85EBCA77u + 0x27D4EB2Fu);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  wgt = l / p;
  return static_cast<float>(h >> 8) * (1.f / 16777216.f) < p;
}

// The records of segment x in use: the first min(count, capacity) of it.
__device__ __forceinline__ uint32_t rec_used(const WaveBuffers& B, int x) {
  return min(B.counters[ctr_rec(x)], B.rec_cap / kQSeg);
}

// After a guided batch, pass 1 over its training records (grid-stride over
// each segment's records in use): the lambda of every record into lam_all,
// their sum into *lam_sum_all (one atomic per work-group).
__global__ void __launch_bounds__(256) k_guide_lambda(WaveBuffers B, float* lam_all, float* lam_sum_all) {
  __shared__ float red[4];
  const uint32_t rs = B.rec_cap / kQSeg;
  float acc = 0.f;
  for (int x = 0; x < kQSeg; ++x) {
    const uint32_t n = rec_used(B, x);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
      const uint32_t r = static_cast<uint32_t>(x) * rs + j;
      const float l = record_lambda(B, r);
      lam_all[r] = l;
      acc += l;
    }
  }
  const float t = block_sumf(acc, red);
  if (threadIdx.x == 0 && t > 0.f) atomicAdd(lam_sum_all, t);
}

// Pass 2: the training batch, at most `cap` records drawn with probability
// p = min(1, lambda / tau), tau = (sum of lambda) / target, and weighted
// lambda / p = max(lambda, tau). The loss sum over the drawn records is then
// an unbiased estimate of the sum over all of them, while every record with
// lambda >= tau (the rare bright paths a guide must learn, such as
// caustics) is kept and records with lambda 0 (no gradient) never are.
// Each thread walks its records twice (count, then write), so a work-group
// reserves its output range with one atomic. Drawn records go to
// pos/nrm/dir/lam; *n_out counts them, *lam_sum adds their weights.
__global__ void __launch_bounds__(256) k_guide_select(WaveBuffers B, const float* lam_all, const float* lam_sum_all,
                                                      float target, uint32_t seed, float* pos, float* nrm, float* dir,
                                                      float* lam, float* lam_sum, uint32_t* n_out, uint32_t cap) {
  __shared__ uint32_t red[5];
  __shared__ float redf[4];
  __shared__ uint32_t base_sh;
  const uint32_t rs = B.rec_cap / kQSeg;
  const float tau = *lam_sum_all / target;
  const uint32_t r0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  uint32_t mine = 0;
  for (int x = 0; x < kQSeg; ++x) {
    const uint32_t n = rec_used(B, x);
    for (uint32_t j = r0; j < n; j += stride) {
      const uint32_t r = static_cast<uint32_t>(x) * rs + j;
      float w;
      if (record_drawn(r, lam_all[r], tau, seed, w)) ++mine;
    }
  }
  uint32_t total;
  const uint32_t before = block_excl_scan(mine, red, &total);
  if (threadIdx.x == 0) base_sh = total > 0u ? atomicAdd(n_out, total) : 0u;
  __syncthreads();
  uint32_t k = base_sh + before;
  float added = 0.f;
  const uint32_t R = B.rec_cap;
  for (int x = 0; x < kQSeg && k < cap; ++x) {
    const uint32_t n = rec_used(B, x);
    for (uint32_t j = r0; j < n && k < cap; j += stride) {
      const uint32_t r = static_cast<uint32_t>(x) * rs + j;
      float w;
      if (!record_drawn(r, lam_all[r], tau, seed, w)) continue;
      const float4 a = B.rec[r], b = B.rec[R + r], c = B.rec[2 * R + r];
      pos[3 * k] = a.x;
      pos[3 * k + 1] = a.y;
      pos[3 * k + 2] = a.z;
      nrm[3 * k] = b.x;
      nrm[3 * k + 1] = b.y;
      nrm[3 * k + 2] = b.z;
      dir[3 * k] = c.x;
      dir[3 * k + 1] = c.y;
      dir[3 * k + 2] = c.z;
      lam[k] = w;
      added += w;
      ++k;
    }
  }
  const float t = block_sumf(added, redf);
  if (threadIdx.x == 0 && t > 0.f) atomicAdd(lam_sum, t);
}

// Rearranges the FP16 row-major weights into MFMA operand tiles.
__global__ void __launch_bounds__(256) k_tile_weights(const uint16_t* w, uint16_t* wt) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= kWTotal) return;
  int off, K, N;
  if (i < kW2) {
    off = kW1; K = kGIn; N = kGHid;
  } else if (i < kW3) {
    off = kW2; K = kGHid; N = kGHid;
  } else if (i < kW4) {
    off = kW3; K = kGHid; N = kGHid;
  } else {
    off = kW4; K = kGHid; N = kGOut;
  }
  const int r = i - off, k = r / N, c = r - k * N;
  wt[off + tiled_index(k, c, K, N)] = w[i];
}

// Forward + loss + backward over n training samples; accumulates the
// gradients of sum_i l_i into grid_g (L x T x 2) and w_g (kWTotal), and the
// summed loss into loss_sum.
// With n_dev non-null the record count is cnt = min(*n_dev, n); the kernel
// trains on slice `slice` of `nslices` equal parts of [0, cnt) and takes
// lam_scale = cnt / *lam_sum_dev (the batch mean of lambda scaled to 1).
__global__ void __launch_bounds__(64) k_guide_train(GuideNet net, int n, const float* pos, const float* nrm,
                                                    const float* dir, const float* lam, float lam_scale,
                                                    const uint32_t* n_dev, const float* lam_sum_dev, int slice,
                                                    int nslices, float* grid_g, float* w_part, float* loss_sum) {
  __shared__ __align__(16) TrainLds s;
  float* w_g = w_part + static_cast<size_t>(blockIdx.x) * kWTotal;   // this work-group's weight gradients
  int begin = 0;
  if (n_dev) {
    const uint32_t cnt = min(*n_dev, static_cast<uint32_t>(n));
    const float sum = *lam_sum_dev;
    lam_scale = sum > 0.f ? static_cast<float>(cnt) / sum : 0.f;
    begin = static_cast<int>((static_cast<uint64_t>(cnt) * slice) / nslices);
    n = static_cast<int>((static_cast<uint64_t>(cnt) * (slice + 1)) / nslices);
  }
  if (begin + static_cast<int>(blockIdx.x) * 64 >= n) return;
  const int lane = glane();
  const int i = begin + static_cast<int>(blockIdx.x) * 64 + lane;
  const bool ok = i < n;
  const f3 p = ok ? f3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]} : f3{net.lo[0], net.lo[1], net.lo[2]};
  const f3 nn = ok ? f3{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]} : f3{0.f, 0.f, 1.f};
  const f3 w = ok ? f3{dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]} : f3{0.f, 0.f, 1.f};
  const float lm = ok ? lam[i] * lam_scale : 0.f;
  guide_encode(net, p, nn, s.X + lane * kGIn);
  __syncthreads();
  mlp_forward(net, s.X, s.H1, s.H2, s.H3, s.F);
  // output gradients -> Ga (64 x 48, FP16)
  float g[kGOut];
  const float l = loss_grad(s.F + lane * kGOut, w, lm, g);
  for (int k = 0; k < kGOut; ++k) s.Ga[lane * kGOut + k] = static_cast<_Float16>(g[k]);
  const float lsum = g_wave_sumf(ok ? l : 0.f);
  if (lane == 0) atomicAdd(loss_sum, lsum);
  __syncthreads();
  // layer 4: dW4 = H3^T G4 ; G3 = G4 W4^T * (H3 > 0)
  weight_grad(s.H3, kGHid, s.Ga, kGOut, w_g + kW4);
  const _Float16* W = reinterpret_cast<const _Float16*>(net.w);
  input_grad(s.Ga, kGOut, W + kW4, kGHid, s.H3, s.Gb, nullptr);
  weight_grad(s.H2, kGHid, s.Gb, kGHid, w_g + kW3);
  input_grad(s.Gb, kGHid, W + kW3, kGHid, s.H2, s.Ga, nullptr);
  weight_grad(s.H1, kGHid, s.Ga, kGHid, w_g + kW2);
  input_grad(s.Ga, kGHid, W + kW2, kGHid, s.H1, s.Gb, nullptr);
  weight_grad(s.X, kGIn, s.Gb, kGHid, w_g + kW1);
  input_grad(s.Gb, kGHid, W + kW1, kGIn, nullptr, nullptr, s.F);
  // hash-grid scatter of the encoding gradient (first 32 inputs)
  if (ok) {
    const f3 q = guide_unit_pos(net, p);
    const uint32_t T = 1u << net.t_log2;
    for (int lv = 0; lv < kGLevels; ++lv) {
      const float g0 = s.F[lane * kGIn + 2 * lv], g1 = s.F[lane * kGIn + 2 * lv + 1];
      if (g0 == 0.f && g1 == 0.f) continue;
      float* gl = grid_g + static_cast<size_t>(lv) * T * 2u;
      for (int c = 0; c < 8; ++c) {
        uint32_t slot;
        float wt;
        grid_corner(net, q, lv, c, slot, wt);
        if (wt == 0.f) continue;
        atomicAdd(&gl[2u * slot], wt * g0);
        atomicAdd(&gl[2u * slot + 1u], wt * g1);
      }
    }
  }
}

// Sums the per-work-group weight gradients of a k_guide_train launch into
// w_g: thread (w, chunk) adds partials of work-groups chunk*kRedChunk ..
// +kRedChunk for weight w, one atomic per (w, chunk). The number of
// work-groups that ran follows from the slice's record count, as in
// k_guide_train.
constexpr int kRedChunk = 64;
__global__ void __launch_bounds__(256) k_wgrad_reduce(const float* w_part, float* w_g, int n, const uint32_t* n_dev,
                                                      int slice, int nslices) {
  const int w = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (w >= kWTotal) return;
  int begin = 0;
  if (n_dev) {
    const uint32_t cnt = min(*n_dev, static_cast<uint32_t>(n));
    begin = static_cast<int>((static_cast<uint64_t>(cnt) * slice) / nslices);
    n = static_cast<int>((static_cast<uint64_t>(cnt) * (slice + 1)) / nslices);
  }
  const int used = n > begin ? (n - begin + 63) / 64 : 0;
  const int b0 = static_cast<int>(blockIdx.y) * kRedChunk, b1 = min(used, b0 + kRedChunk);
  if (b0 >= b1) return;
  float acc = 0.f;
  for (int b = b0; b < b1; ++b) acc += w_part[static_cast<size_t>(b) * kWTotal + w];
  if (acc != 0.f) atomicAdd(&w_g[w], acc);
}

// Adam with bias correction on FP32 masters; refreshes the FP16 copies and
// clears the gradients. grad_scale converts summed gradients to means.
// With n_dev non-null the gradient scale is nslices / min(*n_dev, cap), one
// over the size of a training slice (no step when there are no records).
__global__ void __launch_bounds__(256) k_guide_adam(float* master, uint16_t* half_bits, float* grad, float* m,
                                                    float* v, size_t count, float lr, float b1, float b2, float eps,
                                                    float bc1, float bc2, float grad_scale, const uint32_t* n_dev,
                                                    uint32_t cap, int nslices) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  if (n_dev) {
    const uint32_t cnt = min(*n_dev, cap);
    if (cnt < static_cast<uint32_t>(nslices)) return;
    grad_scale = static_cast<float>(nslices) / static_cast<float>(cnt);
  }
  const float g = grad[i] * grad_scale;
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float p = master[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  master[i] = p;
  reinterpret_cast<_Float16*>(half_bits)[i] = static_cast<_Float16>(p);
  grad[i] = 0.f;
}

__global__ void __launch_bounds__(256) k_to_half(const float* src, uint16_t* dst, size_t count) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) reinterpret_cast<_Float16*>(dst)[i] = static_cast<_Float16>(src[i]);
}

hipError_t launch_guide_eval(hipStream_t s, const GuideNet& net, int n, const float* pos, const float* nrm,
                             float* raw) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_guide_eval, dim3((n + 63) / 64), dim3(64), 0, s, net, n, pos, nrm, raw);
  return hipGetLastError();
}

hipError_t launch_guide_query(hipStream_t s, const GuideNet& net, int n, const float* pos, const float* nrm,
                              const float* u, float* dir, float* pdf) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_guide_query, dim3((n + 63) / 64), dim3(64), 0, s, net, n, pos, nrm, u, dir, pdf);
  return hipGetLastError();
}

int guide_train_blocks(int n, bool device_count, int nslices) {
  const int per = device_count ? (n + nslices - 1) / nslices + 1 : n;
  return (per + 63) / 64;
}

hipError_t launch_guide_train(hipStream_t s, const GuideNet& net, int n, const float* pos, const float* nrm,
                              const float* dir, const float* lam, float lam_scale, const uint32_t* n_dev,
                              const float* lam_sum_dev, int slice, int nslices, float* grid_g, float* w_part,
                              float* w_g, float* loss_sum) {
  if (n <= 0) return hipSuccess;
  const int blocks = guide_train_blocks(n, n_dev != nullptr, nslices);
  hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(64), 0, s, net, n, pos, nrm, dir, lam, lam_scale, n_dev,
                     lam_sum_dev, slice, nslices, grid_g, w_part, loss_sum);
  const dim3 rgrid((kWTotal + 255) / 256, (blocks + kRedChunk - 1) / kRedChunk);
  hipLaunchKernelGGL(k_wgrad_reduce, rgrid, dim3(256), 0, s, w_part, w_g, n, n_dev, slice, nslices);
  return hipGetLastError();
}

hipError_t launch_guide_adam(hipStream_t s, float* master, uint16_t* half_copy, float* grad, float* m, float* v,
                             size_t count, float lr, float b1, float b2, float eps, float bc1, float bc2,
                             float grad_scale, const uint32_t* n_dev, uint32_t cap, int nslices) {
  const size_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_guide_adam, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, master, half_copy, grad, m,
                     v, count, lr, b1, b2, eps, bc1, bc2, grad_scale, n_dev, cap, nslices);
  return hipGetLastError();
}

hipError_t launch_guide_infer(hipStream_t s, const WaveBuffers& B, int qi, int depth, const GuideNet& net, uint4* feat,
                              int grid) {
  // work-groups of kInferWaves waves; `grid` counts waves
  const int blocks = (grid + kInferWaves - 1) / kInferWaves;
  hipLaunchKernelGGL(k_guide_encode, dim3(blocks), dim3(64 * kInferWaves), 0, s, B, qi, depth, net, feat);
  hipLaunchKernelGGL(k_guide_mlp, dim3(blocks), dim3(64 * kInferWaves), 0, s, B, qi, depth, net,
                     static_cast<const uint4*>(feat));
  return hipGetLastError();
}

hipError_t launch_guide_collect(hipStream_t s, const WaveBuffers& B, float* lam_all, float* lam_sum_all,
                                float target, uint32_t seed, float* pos, float* nrm, float* dir, float* lam,
                                float* lam_sum, uint32_t* n_out, uint32_t cap) {
  // grid-stride passes: about 16 records per thread at full capacity
  const uint32_t n = B.rec_cap > 0u ? B.rec_cap : 1u;
  const uint32_t blocks = std::max(1u, std::min(4096u, (n + 4095u) / 4096u));
  hipLaunchKernelGGL(k_guide_lambda, dim3(blocks), dim3(256), 0, s, B, lam_all, lam_sum_all);
  hipLaunchKernelGGL(k_guide_select, dim3(blocks), dim3(256), 0, s, B, lam_all, lam_sum_all, target, seed, pos, nrm,
                     dir, lam, lam_sum, n_out, cap);
  return hipGetLastError();
}

hipError_t launch_tile_weights(hipStream_t s, const uint16_t* w, uint16_t* wt) {
  hipLaunchKernelGGL(k_tile_weights, dim3((kWTotal + 255) / 256), dim3(256), 0, s, w, wt);
  return hipGetLastError();
}

hipError_t launch_to_half(hipStream_t s, const float* src, uint16_t* dst, size_t count) {
  const size_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_to_half, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, src, dst, count);
  return hipGetLastError();
}

}  // namespace xnh
