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

// wave: dW (K x N, global FP32) += A^T (K x 64) * G (64 x N); A: 64 x K rows.
__device__ __forceinline__ void weight_grad(const _Float16* A, int K, const _Float16* G, int N, float* dW) {
  const int lane = glane();
  for (int rt = 0; rt < K / 16; ++rt)
    for (int ct = 0; ct < N / 16; ++ct) {
      const f4v acc = mma_tile<true, false>(A, K, 16 * rt, G, N, 16 * ct, 64);
      for (int j = 0; j < 4; ++j) {
        const int r = 16 * rt + 4 * (lane >> 4) + j, col = 16 * ct + (lane & 15);
        if (acc[j] != 0.f) atomicAdd(&dW[r * N + col], acc[j]);
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
  __shared__ TrainLds s;
  const int lane = glane();
  const int i = static_cast<int>(blockIdx.x) * 64 + lane;
  const bool ok = i < n;
  const f3 p = ok ? f3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]} : f3{net.lo[0], net.lo[1], net.lo[2]};
  const f3 nn = ok ? f3{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]} : f3{0.f, 0.f, 1.f};
  guide_encode(net, p, nn, s.X + lane * kGIn);
  __syncthreads();
  mlp_forward(net, s.X, s.H1, s.H2, s.H3, s.F);
  if (ok)
    for (int k = 0; k < kGOut; ++k) raw[static_cast<size_t>(i) * kGOut + k] = s.F[lane * kGOut + k];
}

// Mixture queries for n points: with u non-null, draws one direction per
// point from its mixture with (u0, u1, u2) = u[3i..3i+2] and writes it to
// dir; then writes the mixture pdf at dir[i] (drawn or given) to pdf.
__global__ void __launch_bounds__(64) k_guide_query(GuideNet net, int n, const float* pos, const float* nrm,
                                                    const float* u, float* dir, float* pdf) {
  __shared__ TrainLds s;
  const int lane = glane();
  const int i = static_cast<int>(blockIdx.x) * 64 + lane;
  const bool ok = i < n;
  const f3 p = ok ? f3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]} : f3{net.lo[0], net.lo[1], net.lo[2]};
  const f3 nn = ok ? f3{nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]} : f3{0.f, 0.f, 1.f};
  guide_encode(net, p, nn, s.X + lane * kGIn);
  __syncthreads();
  mlp_forward(net, s.X, s.H1, s.H2, s.H3, s.F);
  if (!ok) return;
  Mixture m;
  decode_mixture(s.F + lane * kGOut, m);
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

// Forward + loss + backward over n training samples; accumulates the
// gradients of sum_i l_i into grid_g (L x T x 2) and w_g (kWTotal), and the
// summed loss into loss_sum.
__global__ void __launch_bounds__(64) k_guide_train(GuideNet net, int n, const float* pos, const float* nrm,
                                                    const float* dir, const float* lam, float lam_scale,
                                                    float* grid_g, float* w_g, float* loss_sum) {
  __shared__ TrainLds s;
  const int lane = glane();
  const int i = static_cast<int>(blockIdx.x) * 64 + lane;
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

// Adam with bias correction on FP32 masters; refreshes the FP16 copies and
// clears the gradients. grad_scale converts summed gradients to means.
__global__ void __launch_bounds__(256) k_guide_adam(float* master, uint16_t* half_bits, float* grad, float* m,
                                                    float* v, size_t count, float lr, float b1, float b2, float eps,
                                                    float bc1, float bc2, float grad_scale) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
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

hipError_t launch_guide_train(hipStream_t s, const GuideNet& net, int n, const float* pos, const float* nrm,
                              const float* dir, const float* lam, float lam_scale, float* grid_g, float* w_g,
                              float* loss_sum) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_guide_train, dim3((n + 63) / 64), dim3(64), 0, s, net, n, pos, nrm, dir, lam, lam_scale,
                     grid_g, w_g, loss_sum);
  return hipGetLastError();
}

hipError_t launch_guide_adam(hipStream_t s, float* master, uint16_t* half_copy, float* grad, float* m, float* v,
                             size_t count, float lr, float b1, float b2, float eps, float bc1, float bc2,
                             float grad_scale) {
  const size_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_guide_adam, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, master, half_copy, grad, m,
                     v, count, lr, b1, b2, eps, bc1, bc2, grad_scale);
  return hipGetLastError();
}

hipError_t launch_to_half(hipStream_t s, const float* src, uint16_t* dst, size_t count) {
  const size_t blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_to_half, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, src, dst, count);
  return hipGetLastError();
}

}  // namespace xnh
