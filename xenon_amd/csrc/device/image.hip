// image.hip — the output stage on the device, where the image already is:
// the 8-bit conversion the PNG writer applies (main.cpp:25-31) and the
// reference-free noise estimate NoiseStats (math/image.h:35-182).
//
// Every per-pixel value is the reference's float expression in its order
// (built with -ffp-contract=off, so nothing is fused): luminance, the 1-2-1
// blur of interior pixels (which never reach the clamped border), the
// gradient edge test and the two normalised residuals are bit-identical.
// The reference sums luminance and residuals in double in row order; here
// the sums are double too, in a tree order, so the float results can differ
// only where a double sum lands within its own rounding error of a float
// rounding boundary. The median and 95th percentile are selected exactly:
// the residuals are non-negative floats, whose bit patterns order like the
// values, and a four-pass radix select (8 bits a pass, both ranks at once)
// finds the k-th smallest pattern with k computed as the reference computes
// it. Nothing leaves the device until the 64-byte result.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <mutex>

#include "launch.h"

namespace xnh {
namespace {

constexpr int kImgThreads = 256;
constexpr uint32_t kNoKey = 0xFFFFFFFFu;   // above every non-negative float's bit pattern

__device__ __forceinline__ double wave_sumd(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ uint32_t wave_sumu(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// image.h:94-95: the running maximum starts at 0 and takes a value only if
// it compares greater (a NaN never does); combining partial maxima by the
// same rule gives the same result
__device__ __forceinline__ float gt_max(float m, float v) { return v > m ? v : m; }

__device__ __forceinline__ float wave_gt_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = gt_max(v, __shfl_xor(v, o));
  return v;
}

// std::max(a, b) returns b only when a < b
__device__ __forceinline__ float std_max(float a, float b) { return a < b ? b : a; }

// Block sums (kImgThreads threads); red holds one slot per wave.
__device__ double block_sumd(double v, double* red) {
  v = wave_sumd(v);
  const int w = static_cast<int>(threadIdx.x >> 6), nw = static_cast<int>(blockDim.x >> 6);
  if ((threadIdx.x & 63u) == 0u) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < nw; ++k) t += red[k];
  __syncthreads();
  return t;
}

__global__ void __launch_bounds__(kImgThreads) k_img_from_accum(const float4* accum, float inv_spp, int w, int h,
                                                                float* img) {
  // the swapchain layout (cpu_tiled.cpp:224-234): pixel row y goes to row h-1-y
  const int p = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= w * h) return;
  const int y = p / w, x = p - y * w;
  const float4 a = accum[p];
  float* o = img + 3 * (static_cast<size_t>(h - 1 - y) * w + x);
  o[0] = a.x * inv_spp;
  o[1] = a.y * inv_spp;
  o[2] = a.z * inv_spp;
}

// image.h:127-135 and 137-146: luminance, and per-block double sums of it.
__global__ void __launch_bounds__(kImgThreads) k_img_lum(const float* img, int n, float* lum, double* part) {
  __shared__ double red[kImgThreads / 64];
  double acc = 0.0;
  for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < n;
       i += static_cast<int>(gridDim.x * blockDim.x)) {
    const float r = img[3 * static_cast<size_t>(i)], g = img[3 * static_cast<size_t>(i) + 1],
                b = img[3 * static_cast<size_t>(i) + 2];
    const float l = 0.2126f * r + 0.7152f * g + 0.0722f * b;
    lum[i] = l;
    acc += static_cast<double>(l);
  }
  const double t = block_sumd(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// image.h:51-57: average luminance, edge threshold and epsilon.
__global__ void __launch_bounds__(kImgThreads) k_img_prep(const double* part, int nb, int n, ImgState* st) {
  __shared__ double red[kImgThreads / 64];
  double acc = 0.0;
  for (int k = static_cast<int>(threadIdx.x); k < nb; k += static_cast<int>(blockDim.x)) acc += part[k];
  const double t = block_sumd(acc, red);
  if (threadIdx.x == 0) {
    const float avg = static_cast<float>(t / static_cast<double>(n));
    st->avg = avg;
    st->edge = std_max(1e-6f, avg * 0.1f);
    st->eps = std_max(1e-8f, avg * avg * 1e-4f);
  }
}

// image.h:63-98 for every interior pixel: the edge test, the residual
// against the blur and its two normalised forms. key[i] is the absolute
// form's bit pattern for a counted pixel, kNoKey otherwise; per block the
// double sums of both forms, the maximum and the count.
__global__ void __launch_bounds__(kImgThreads) k_img_resid(const float* lum, int w, int h, const ImgState* st,
                                                           uint32_t* key, double* part2, float* pmax,
                                                           uint32_t* pcnt) {
  __shared__ double red[kImgThreads / 64];
  __shared__ float redm[kImgThreads / 64];
  __shared__ uint32_t redc[kImgThreads / 64];
  const float edge = st->edge, eps = st->eps;
  const int n = w * h;
  double s2 = 0.0, s1 = 0.0;
  float mx = 0.f;
  uint32_t cnt = 0u;
  for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < n;
       i += static_cast<int>(gridDim.x * blockDim.x)) {
    const int y = i / w, x = i - y * w;
    uint32_t k = kNoKey;
    if (x >= 1 && x < w - 1 && y >= 1 && y < h - 1) {
      const float gx = 0.5f * (lum[i + 1] - lum[i - 1]);
      const float gy = 0.5f * (lum[i + w] - lum[i - w]);
      const float grad = sqrtf(gx * gx + gy * gy);
      if (!(grad >= edge)) {
        // image.h:152-174 at an interior pixel: weights 1 2 1 / 2 4 2 / 1 2 1
        // summed row by row, left to right, then divided by 16
        float s = 0.f;
        s += 1.f * lum[i - w - 1];
        s += 2.f * lum[i - w];
        s += 1.f * lum[i - w + 1];
        s += 2.f * lum[i - 1];
        s += 4.f * lum[i];
        s += 2.f * lum[i + 1];
        s += 1.f * lum[i + w - 1];
        s += 2.f * lum[i + w];
        s += 1.f * lum[i + w + 1];
        const float bl = s / 16.f;
        const float r = lum[i] - bl;
        const float den = bl * bl + eps;
        const float q2 = (r * r) / den;
        const float q1 = fabsf(r) / sqrtf(den);
        s2 += static_cast<double>(q2);
        s1 += static_cast<double>(q1);
        mx = gt_max(mx, q1);
        ++cnt;
        k = __float_as_uint(q1);
      }
    }
    key[i] = k;
  }
  const double t2 = block_sumd(s2, red);
  const double t1 = block_sumd(s1, red);
  const float wm = wave_gt_max(mx);
  const uint32_t wc = wave_sumu(cnt);
  const int wv = static_cast<int>(threadIdx.x >> 6), nw = static_cast<int>(blockDim.x >> 6);
  if ((threadIdx.x & 63u) == 0u) {
    redm[wv] = wm;
    redc[wv] = wc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    uint32_t c = 0u;
    for (int k = 0; k < nw; ++k) {
      m = gt_max(m, redm[k]);
      c += redc[k];
    }
    part2[2 * blockIdx.x] = t2;
    part2[2 * blockIdx.x + 1] = t1;
    pmax[blockIdx.x] = m;
    pcnt[blockIdx.x] = c;
  }
}

// image.h:100-118: totals, the three means, and the two ranks
// k = (size_t)(p * (float)(n - 1)) of image.h:176-186.
__global__ void __launch_bounds__(kImgThreads) k_img_final(const double* part2, const float* pmax,
                                                           const uint32_t* pcnt, int nb, ImgState* st) {
  __shared__ double red[kImgThreads / 64];
  __shared__ float redm[kImgThreads / 64];
  __shared__ uint32_t redc[kImgThreads / 64];
  double s2 = 0.0, s1 = 0.0;
  float mx = 0.f;
  uint32_t cnt = 0u;
  for (int k = static_cast<int>(threadIdx.x); k < nb; k += static_cast<int>(blockDim.x)) {
    s2 += part2[2 * k];
    s1 += part2[2 * k + 1];
    mx = gt_max(mx, pmax[k]);
    cnt += pcnt[k];
  }
  const double t2 = block_sumd(s2, red);
  const double t1 = block_sumd(s1, red);
  const float wm = wave_gt_max(mx);
  const uint32_t wc = wave_sumu(cnt);
  const int wv = static_cast<int>(threadIdx.x >> 6), nw = static_cast<int>(blockDim.x >> 6);
  if ((threadIdx.x & 63u) == 0u) {
    redm[wv] = wm;
    redc[wv] = wc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    uint32_t c = 0u;
    for (int k = 0; k < nw; ++k) {
      m = gt_max(m, redm[k]);
      c += redc[k];
    }
    st->count = c;
    st->max_abs = m;
    st->mean_sq = 0.f;
    st->mean_abs = 0.f;
    st->rank[0] = st->rank[1] = 0u;
    if (c > 0u) {
      st->mean_sq = static_cast<float>(t2 / static_cast<double>(c));
      st->mean_abs = static_cast<float>(t1 / static_cast<double>(c));
      st->rank[0] = static_cast<uint32_t>(0.50f * static_cast<float>(c - 1u));
      st->rank[1] = static_cast<uint32_t>(0.95f * static_cast<float>(c - 1u));
    }
    st->prefix[0] = st->prefix[1] = 0u;
  }
  for (int k = static_cast<int>(threadIdx.x); k < 512; k += static_cast<int>(blockDim.x)) st->hist[k] = 0u;
}

// One radix-select pass: histogram of digit `pass` (most significant first)
// of the keys that carry each target's prefix so far.
__global__ void __launch_bounds__(kImgThreads) k_img_hist(const uint32_t* key, int n, ImgState* st, int pass) {
  __shared__ uint32_t lh[512];
  for (int k = static_cast<int>(threadIdx.x); k < 512; k += static_cast<int>(blockDim.x)) lh[k] = 0u;
  __syncthreads();
  const int shift = 24 - 8 * pass;
  const uint32_t p0 = st->prefix[0], p1 = st->prefix[1];
  for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < n;
       i += static_cast<int>(gridDim.x * blockDim.x)) {
    const uint32_t k = key[i];
    if (k == kNoKey) continue;
    const uint32_t d = (k >> shift) & 255u;
    const uint32_t hi = pass == 0 ? 0u : (k >> (shift + 8));
    if (hi == p0) atomicAdd(&lh[d], 1u);
    if (hi == p1) atomicAdd(&lh[256 + d], 1u);
  }
  __syncthreads();
  for (int k = static_cast<int>(threadIdx.x); k < 512; k += static_cast<int>(blockDim.x))
    if (lh[k]) atomicAdd(&st->hist[k], lh[k]);
}

// The digit whose bin holds each target's rank; the rank becomes the rank
// within that bin, and the histogram is cleared for the next pass.
__global__ void __launch_bounds__(64) k_img_pick(ImgState* st) {
  const int t = static_cast<int>(threadIdx.x);
  if (t < 2 && st->count > 0u) {
    uint32_t cum = 0u, r = st->rank[t];
    for (uint32_t d = 0; d < 256u; ++d) {
      const uint32_t c = st->hist[256 * t + d];
      if (r < cum + c) {
        st->prefix[t] = (st->prefix[t] << 8) | d;
        st->rank[t] = r - cum;
        break;
      }
      cum += c;
    }
  }
  __syncthreads();
  for (int k = t; k < 512; k += static_cast<int>(blockDim.x)) st->hist[k] = 0u;
}

// main.cpp:28-30 with the writer's vertical flip: file row r is framebuffer
// row h-1-r. powf is taken as the double power rounded to float, which is
// the correctly rounded powf; glibc's powf differs from it on 671,164 of the
// 1,065,353,217 floats in [0, 1] and none of them changes the byte.
__global__ void __launch_bounds__(kImgThreads) k_img_rgb8(const float* img, int w, int h, uint8_t* out) {
  const int q = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= 3 * w * h) return;
  const int r = q / (3 * w), k = q - r * 3 * w;
  float v = img[3 * static_cast<size_t>(h - 1 - r) * w + k];
  v = v < 0.f ? 0.f : (1.f < v ? 1.f : v);   // std::clamp
  const float c = static_cast<float>(pow(static_cast<double>(v), static_cast<double>(1.f / 2.2f)));
  out[q] = static_cast<uint8_t>(static_cast<uint32_t>(255.99f * c));
}

int img_blocks(int n) {
  const int b = (n + kImgThreads - 1) / kImgThreads;
  return b < 1 ? 1 : (b > kImgMaxBlocks ? kImgMaxBlocks : b);
}

}  // namespace

hipError_t launch_img_from_accum(hipStream_t s, const float4* accum, float inv_spp, int w, int h, float* img) {
  const int n = w * h;
  hipLaunchKernelGGL(k_img_from_accum, dim3((n + kImgThreads - 1) / kImgThreads), dim3(kImgThreads), 0, s, accum,
                     inv_spp, w, h, img);
  return hipGetLastError();
}

hipError_t launch_noise_stats(hipStream_t s, const float* img, int w, int h, ImgScratch sc) {
  const int n = w * h, nb = img_blocks(n);
  hipLaunchKernelGGL(k_img_lum, dim3(nb), dim3(kImgThreads), 0, s, img, n, sc.lum, sc.part);
  hipLaunchKernelGGL(k_img_prep, dim3(1), dim3(kImgThreads), 0, s, sc.part, nb, n, sc.st);
  hipLaunchKernelGGL(k_img_resid, dim3(nb), dim3(kImgThreads), 0, s, sc.lum, w, h, sc.st, sc.key, sc.part,
                     sc.pmax, sc.pcnt);
  hipLaunchKernelGGL(k_img_final, dim3(1), dim3(kImgThreads), 0, s, sc.part, sc.pmax, sc.pcnt, nb, sc.st);
  for (int pass = 0; pass < 4; ++pass) {
    hipLaunchKernelGGL(k_img_hist, dim3(nb), dim3(kImgThreads), 0, s, sc.key, n, sc.st, pass);
    hipLaunchKernelGGL(k_img_pick, dim3(1), dim3(64), 0, s, sc.st);
  }
  return hipGetLastError();
}

// ---- scratch of the allocation-free NoiseStats entry point ----------------
// One static device array per GPU, carved into the ImgScratch fields for
// images of at most kImgStaticMaxPixels. The launch sequence above writes
// every scratch word it later reads within the same call (the partial sums
// per block, the keys per pixel; k_img_final clears the histogram and the
// radix-select state), so nothing needs clearing between calls. Calls that
// overlap on one GPU share the array and race.
namespace {
constexpr size_t align256(size_t x) { return (x + 255u) & ~static_cast<size_t>(255u); }
constexpr size_t kStOffLum = 0;
constexpr size_t kStOffKey = align256(kStOffLum + sizeof(float) * kImgStaticMaxPixels);
constexpr size_t kStOffPart = align256(kStOffKey + sizeof(uint32_t) * kImgStaticMaxPixels);
constexpr size_t kStOffPmax = align256(kStOffPart + sizeof(double) * 2 * kImgMaxBlocks);
constexpr size_t kStOffPcnt = align256(kStOffPmax + sizeof(float) * kImgMaxBlocks);
constexpr size_t kStOffState = align256(kStOffPcnt + sizeof(uint32_t) * kImgMaxBlocks);
constexpr size_t kStBytes = align256(kStOffState + sizeof(ImgState));
static_assert(kStBytes == kImgStaticScratchBytes, "static scratch size");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};   // the array's device address, per device
}  // namespace

static __device__ __align__(256) unsigned char scratch[kStBytes];

hipError_t img_static_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device) return hipErrorInvalidDevice;   // the address is the current device's
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch[device] = static_cast<unsigned char*>(p);
  return hipSuccess;
}

hipError_t img_static_scratch(int device, ImgScratch* sc) {
  if (device < 0 || device >= kMaxDevices || !sc) return hipErrorInvalidDevice;
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = g_scratch[device];
  }
  if (!base) return hipErrorNotInitialized;
  sc->lum = reinterpret_cast<float*>(base + kStOffLum);
  sc->key = reinterpret_cast<uint32_t*>(base + kStOffKey);
  sc->part = reinterpret_cast<double*>(base + kStOffPart);
  sc->pmax = reinterpret_cast<float*>(base + kStOffPmax);
  sc->pcnt = reinterpret_cast<uint32_t*>(base + kStOffPcnt);
  sc->st = reinterpret_cast<ImgState*>(base + kStOffState);
  return hipSuccess;
}

hipError_t launch_img_rgb8(hipStream_t s, const float* img, int w, int h, uint8_t* out) {
  const int q = 3 * w * h;
  hipLaunchKernelGGL(k_img_rgb8, dim3((q + kImgThreads - 1) / kImgThreads), dim3(kImgThreads), 0, s, img, w, h, out);
  return hipGetLastError();
}

}  // namespace xnh
