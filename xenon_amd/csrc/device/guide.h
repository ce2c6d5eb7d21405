// guide.h — the path-guiding network on the device: a multiresolution hash
// grid and a bias-free 3x64 ReLU MLP mapping (position, normal) to an
// 8-lobe von Mises-Fisher mixture over incident direction.
//
// Methods: hash encoding of Mueller et al. 2022 (levels, geometric
// resolutions, spatial hash x*1 ^ y*2654435761 ^ z*805459861, dense tables
// where a level fits, trilinear weights); real spherical harmonics of degree
// <= 3 for the normal; vMF lobes with C(k) = k / (2 pi (1 - e^{-2k})) and
// Jakob's 2012 stable inverse-CDF sampler; softmax lobe weights. No
// counterpart in the reference renderer.
//
// Layout: grid params FP16 [level][slot][2]; MLP weights FP16 row-major
// [in][out] per layer (48x64, 64x64, 64x64, 64x48). Matrix products use
// v_mfma_f32_16x16x32_f16, and v_mfma_f32_16x16x16f16 for a last 16-wide
// step (FP16 operands, FP32 accumulation), one wave per 64 samples; operand
// lane layout (checked on gfx950): lane l holds D[4(l/16)+j][l%16], j =
// 0..3, and its A and B values pair by position (16x16x16: A[l%16][4(l/16)+j],
// B[4(l/16)+j][l%16]; 16x16x32: 8 values each, any order common to A and B).
// Functions marked "wave" must be reached by all 64 lanes of a one-wave
// block (they use __syncthreads as the LDS barrier).
#pragma once
#include <hip/hip_runtime.h>

#include "../common/guide_net.h"
#include "dmath.h"

namespace xnh {

typedef _Float16 h4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr float kKappaLogMin = -4.0f, kKappaLogMax = 8.0f;
constexpr float kPosMax = 0.99999994f;   // largest float below 1
constexpr float kGuidePi = 3.14159265358979323846f;

struct Mixture {
  f3 mu[kGLobes];
  float kappa[kGLobes];
  float pi[kGLobes];
};

// ---- encoding ---------------------------------------------------------------

__device__ __forceinline__ uint32_t grid_slot(int ix, int iy, int iz, int res, int t_log2) {
  const uint32_t n = static_cast<uint32_t>(res) + 1u;
  const uint64_t n3 = static_cast<uint64_t>(n) * n * n;
  if (n3 <= (1ull << t_log2)) return static_cast<uint32_t>(ix) + static_cast<uint32_t>(iy) * n +
                                     static_cast<uint32_t>(iz) * n * n;
  const uint32_t h = (static_cast<uint32_t>(ix) * 1u) ^ (static_cast<uint32_t>(iy) * 2654435761u) ^
                     (static_cast<uint32_t>(iz) * 805459861u);
  return h & ((1u << t_log2) - 1u);
}

__device__ __forceinline__ f3 guide_unit_pos(const GuideNet& net, f3 p) {
  f3 q{(p.x - net.lo[0]) * net.inv_ext[0], (p.y - net.lo[1]) * net.inv_ext[1], (p.z - net.lo[2]) * net.inv_ext[2]};
  q.x = fminf(fmaxf(q.x, 0.f), kPosMax);
  q.y = fminf(fmaxf(q.y, 0.f), kPosMax);
  q.z = fminf(fmaxf(q.z, 0.f), kPosMax);
  return q;
}

// Trilinear corner of level `l`: slot and weight of corner c (0..7).
__device__ __forceinline__ void grid_corner(const GuideNet& net, f3 q, int l, int c, uint32_t& slot, float& w) {
  const float fr = static_cast<float>(net.res[l]);
  const float x = q.x * fr, y = q.y * fr, z = q.z * fr;
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  const float ax = x - fx, ay = y - fy, az = z - fz;
  const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
  w = (dx ? ax : 1.f - ax) * (dy ? ay : 1.f - ay) * (dz ? az : 1.f - az);
  slot = grid_slot(static_cast<int>(fx) + dx, static_cast<int>(fy) + dy, static_cast<int>(fz) + dz, net.res[l],
                   net.t_log2);
}

// Real SH, degrees 0..3, of a unit vector.
__device__ __forceinline__ void sh16(f3 n, float* o) {
  const float x = n.x, y = n.y, z = n.z;
  const float xx = x * x, yy = y * y, zz = z * z;
  o[0] = 0.28209479177387814f;
  o[1] = -0.48860251190291987f * y;
  o[2] = 0.48860251190291987f * z;
  o[3] = -0.48860251190291987f * x;
  o[4] = 1.0925484305920792f * x * y;
  o[5] = -1.0925484305920792f * y * z;
  o[6] = 0.94617469575755997f * zz - 0.31539156525251999f;
  o[7] = -1.0925484305920792f * x * z;
  o[8] = 0.54627421529603959f * (xx - yy);
  o[9] = 0.59004358992664352f * y * (-3.0f * xx + yy);
  o[10] = 2.8906114426405538f * x * y * z;
  o[11] = 0.45704579946446572f * y * (1.0f - 5.0f * zz);
  o[12] = 0.3731763325901154f * z * (5.0f * zz - 3.0f);
  o[13] = 0.45704579946446572f * x * (1.0f - 5.0f * zz);
  o[14] = 1.4453057213202769f * z * (xx - yy);
  o[15] = 0.59004358992664352f * x * (-xx + 3.0f * yy);
}

// The 48 FP16 encoding values of (p, n), packed two per word (the lower
// address in the low half): words 0..15 the 16 levels' two features, 16..23
// the 16 SH coefficients. The corner fetches of four levels at a time are
// issued before any is used, so a wave waits on memory four times per point
// rather than sixteen; each level still sums its corners in order 0..7.
__device__ __forceinline__ uint32_t pack_h2(float a, float b) {
  const _Float16 ha = static_cast<_Float16>(a), hb = static_cast<_Float16>(b);
  uint16_t ua, ub;
  __builtin_memcpy(&ua, &ha, 2);
  __builtin_memcpy(&ub, &hb, 2);
  return static_cast<uint32_t>(ua) | (static_cast<uint32_t>(ub) << 16);
}

__device__ __forceinline__ void guide_encode_packed(const GuideNet& net, f3 p, f3 n, uint32_t (&o)[24]) {
  const f3 q = guide_unit_pos(net, p);
  const size_t T = static_cast<size_t>(1) << net.t_log2;
  const uint32_t* g32 = reinterpret_cast<const uint32_t*>(net.grid);   // the two features of a slot
  for (int l0 = 0; l0 < kGLevels; l0 += 4) {
    uint32_t pr[32];
    float w[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      const int l = l0 + (k >> 3);
      uint32_t s;
      grid_corner(net, q, l, k & 7, s, w[k]);
      pr[k] = gld1(g32, static_cast<size_t>(l) * T + s);
    }
#pragma unroll
    for (int lk = 0; lk < 4; ++lk) {
      float a0 = 0.f, a1 = 0.f;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const uint32_t pair = pr[8 * lk + c];
        _Float16 h0, h1;
        __builtin_memcpy(&h0, &pair, 2);
        __builtin_memcpy(&h1, reinterpret_cast<const char*>(&pair) + 2, 2);
        a0 = a0 + w[8 * lk + c] * static_cast<float>(h0);
        a1 = a1 + w[8 * lk + c] * static_cast<float>(h1);
      }
      o[l0 + lk] = pack_h2(a0, a1);
    }
  }
  float s[16];
  sh16(n, s);
#pragma unroll
  for (int k = 0; k < 8; ++k) o[16 + k] = pack_h2(s[2 * k], s[2 * k + 1]);
}

// The same 48 values written to `row` (kGIn FP16).
__device__ __forceinline__ void guide_encode(const GuideNet& net, f3 p, f3 n, _Float16* row) {
  uint32_t o[24];
  guide_encode_packed(net, p, n, o);
  uint32_t* r32 = reinterpret_cast<uint32_t*>(row);
#pragma unroll
  for (int k = 0; k < 24; ++k) r32[k] = o[k];
}

// LDS written by some lanes of a wave and read by others: the wave's LDS
// operations complete in order, so a compiler fence at wavefront scope is
// enough (no work-group barrier, which would need every wave present).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- MFMA tiles -----------------------------------------------------------------

// 16x16 tile of op(A) * op(B) over K (a multiple of 16). A: 64 x K rows
// (row-major, lda), rows r0..r0+15 (or its transpose when TA: element (m, k)
// read at A[k * lda + m]); B: K x N row-major (ldb), cols c0..c0+15 (or its
// transpose when TB: element (k, n) read at B[n * ldb + k]). Steps of 32
// inputs use gfx950's 16x16x32 form with lane l holding contraction index
// k0 + 8 (l/16) + j, j = 0..7, in both operands, so an operand contiguous
// along k (A rows, B transposed) is one 16-byte load per lane; a final
// 16-input step (K % 32 = 16) uses the 16x16x16 form. Operands contiguous
// along k must be 16-byte aligned with lda / ldb multiples of 8.
template <bool TA, bool TB>
__device__ __forceinline__ f4v mma_tile(const _Float16* A, int lda, int r0, const _Float16* B, int ldb, int c0,
                                        int K) {
  typedef _Float16 h8 __attribute__((ext_vector_type(8)));
  const int lane = static_cast<int>(threadIdx.x & 63u);
  const int li = lane & 15, g = lane >> 4;
  f4v acc = {0.f, 0.f, 0.f, 0.f};
  int k0 = 0;
  for (; k0 + 32 <= K; k0 += 32) {
    const int kb = k0 + 8 * g;
    h8 a, b;
    if (TA) {
      for (int j = 0; j < 8; ++j) a[j] = A[(kb + j) * lda + r0 + li];
    } else {
      a = *reinterpret_cast<const h8*>(A + (r0 + li) * lda + kb);
    }
    if (TB) {
      b = *reinterpret_cast<const h8*>(B + (c0 + li) * ldb + kb);
    } else {
      for (int j = 0; j < 8; ++j) b[j] = B[(kb + j) * ldb + c0 + li];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
  }
  // The 16-input steps accumulate apart and are added at the end: chaining
  // one MFMA form's result into the other form's input C gave wrong sums on
  // gfx950 (the compiler did not separate the two by enough wait states).
  f4v acc16 = {0.f, 0.f, 0.f, 0.f};
  for (; k0 < K; k0 += 16) {
    const int kb = k0 + 4 * g;
    h4v a, b;
    for (int j = 0; j < 4; ++j) {
      a[j] = TA ? A[(kb + j) * lda + r0 + li] : A[(r0 + li) * lda + kb + j];
      b[j] = TB ? B[(c0 + li) * ldb + kb + j] : B[(kb + j) * ldb + c0 + li];
    }
    acc16 = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, acc16, 0, 0, 0);
  }
  return acc + acc16;
}

// wave: Y (64 x N) = act(X (64 x K) W (K x N)); X, Y in LDS, W global.
// relu: store FP16 ReLU output; else store FP32 (Y32).
__device__ __forceinline__ void dense_fwd(const _Float16* X, int K, const _Float16* W, int N, _Float16* Y, float* Y32) {
  const int lane = static_cast<int>(threadIdx.x & 63u);
  for (int rt = 0; rt < 4; ++rt)
    for (int ct = 0; ct < N / 16; ++ct) {
      const f4v acc = mma_tile<false, false>(X, K, 16 * rt, W, N, 16 * ct, K);
      for (int j = 0; j < 4; ++j) {
        const int r = 16 * rt + 4 * (lane >> 4) + j, col = 16 * ct + (lane & 15);
        if (Y) Y[r * N + col] = static_cast<_Float16>(fmaxf(acc[j], 0.f));
        else Y32[r * N + col] = acc[j];
      }
    }
  __syncthreads();
}

// ---- register-resident forward pass -----------------------------------------------
//
// The MLP evaluated transposed, Y^T = W^T X^T, on gfx950's
// v_mfma_f32_16x16x32_f16 (twice the work of the 16x16x16 form per
// instruction): A = 16 outputs x 32 inputs of W^T, B = 32 features x 16
// samples of X^T, lane l holding 8 values of each, D as for 16x16x16 (lane
// l: output 4 (l/16) + j of sample l % 16, j = 0..3). The accumulators of
// two adjacent output tiles of one layer, concatenated, are the B operand
// of one 32-input step of the next layer: lane l then holds the 32-block's
// features 4 (l/16) + j and 16 + 4 (l/16) + j, j = 0..3, and the weights
// are tiled in that same order, so hidden activations never leave registers
// (any consistent order of the contraction index gives the same product).
// An input width that is not a multiple of 32 (the 48 encoded inputs) ends
// with one 16x16x16 step. Weight tiles (GuideNet::wt), layer L (K x N):
// 32-input tile (mt, kt) at L's offset + (mt * (K/32) + kt) * 512, lane l's
// eight values at + 8 l + j: W[32 kt + f(l/16, j)][16 mt + l % 16] with
// f(g, j) = 4 g + j (j < 4), 16 + 4 g + j - 4 (j >= 4); then, if K % 32 = 16,
// the 16-input tiles mt at + (N/16) (K/32) 512 + 256 mt, lane l's four
// values W[32 (K/32) + 4 (l/16) + j][16 mt + l % 16] at + 4 l + j.
__device__ __forceinline__ int tiled_index(int k, int n, int K, int N) {
  const int nkb = K >> 5, mt = n >> 4, li = n & 15;
  if (k < (nkb << 5)) {
    const int kt = k >> 5, kk = k & 31;
    const int g = (kk & 15) >> 2, j = (kk & 3) + ((kk >> 4) << 2);
    return ((mt * nkb + kt) * 64 + g * 16 + li) * 8 + j;
  }
  const int kk = k - (nkb << 5);
  return (N >> 4) * nkb * 512 + (mt * 64 + (kk >> 2) * 16 + li) * 4 + (kk & 3);
}

typedef _Float16 h8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ h8v cat8(h4v lo, h4v hi) {
  return h8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// One layer of MT output tiles over KT 16-feature input tiles x (KT even,
// or KT = 3 with a 16-input step last); w: the layer's tiles in LDS as h4v
// (16-byte aligned); relu: FP16 ReLU outputs, else the raw FP32 accumulators.
template <int KT, int MT, bool kRelu>
__device__ __forceinline__ void dense_reg(const h4v* x, const h4v* w, h4v* y, f4v* raw = nullptr) {
  const int lane = static_cast<int>(threadIdx.x & 63u);
  constexpr int NKB = KT / 2;
  const h8v* w8 = reinterpret_cast<const h8v*>(w);
  for (int mt = 0; mt < MT; ++mt) {
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < NKB; ++kb)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(w8[(mt * NKB + kb) * 64 + lane], cat8(x[2 * kb], x[2 * kb + 1]), acc, 0,
                                                   0, 0);
    if (KT & 1) {
      // its own accumulator: one MFMA form never takes the other's result
      // as its input C (see mma_tile)
      const f4v zero = {0.f, 0.f, 0.f, 0.f};
      acc += __builtin_amdgcn_mfma_f32_16x16x16f16(w[MT * NKB * 128 + mt * 64 + lane], x[KT - 1], zero, 0, 0, 0);
    }
    if (kRelu) {
      h4v r;
      for (int j = 0; j < 4; ++j) r[j] = static_cast<_Float16>(fmaxf(acc[j], 0.f));
      y[mt] = r;
    } else {
      raw[mt] = acc;
    }
  }
}

// Copies the tiled FP16 weights (GuideNet::wt, kWTotal values) into LDS;
// every thread of the block must call it.
constexpr int kWTiles = kWTotal / 4;   // h4v operands
__device__ __forceinline__ void stage_weights(const uint16_t* wt, h4v* wl) {
  typedef uint16_t u4v __attribute__((ext_vector_type(4)));
  const u4v* src = reinterpret_cast<const u4v*>(wt);
  for (int k = static_cast<int>(threadIdx.x); k < kWTiles; k += static_cast<int>(blockDim.x)) {
    const u4v v = src[k];
    h4v h;
    __builtin_memcpy(&h, &v, sizeof h);
    wl[k] = h;
  }
  __syncthreads();
}

// wave: the MLP on the 64 encoded rows in act (64 x kGIn FP16 LDS rows of
// this wave, written by all its lanes); wl: the weights in LDS
// (stage_weights). On return the lane's 48 raw outputs (FP32 accumulation,
// stored FP16) are at act + lane kGIn.
__device__ __forceinline__ void guide_mlp_wave(_Float16* act, const h4v* wl) {
  const int lane = static_cast<int>(threadIdx.x & 63u), g = lane >> 4, li = lane & 15;
  wave_lds_sync();
  h4v x[4][3];
  for (int nt = 0; nt < 4; ++nt)
    for (int kt = 0; kt < 3; ++kt)
      x[nt][kt] = *reinterpret_cast<const h4v*>(act + (16 * nt + li) * kGIn + 16 * kt + 4 * g);
  wave_lds_sync();   // every row is read before outputs overwrite them
  for (int nt = 0; nt < 4; ++nt) {
    h4v a[4], b[4];
    f4v out[kGOut / 16];
    dense_reg<3, 4, true>(x[nt], wl + kW1 / 4, a);
    dense_reg<4, 4, true>(a, wl + kW2 / 4, b);
    dense_reg<4, 4, true>(b, wl + kW3 / 4, a);
    dense_reg<4, kGOut / 16, false>(a, wl + kW4 / 4, nullptr, out);
    for (int mt = 0; mt < kGOut / 16; ++mt)
      for (int j = 0; j < 4; ++j) act[(16 * nt + li) * kGIn + 16 * mt + 4 * g + j] = static_cast<_Float16>(out[mt][j]);
  }
  wave_lds_sync();
}

// wave: forward pass of 64 points, one per lane (p: world position, n:
// normal); act and wl as for guide_mlp_wave.
__device__ __forceinline__ void guide_forward_wave(const GuideNet& net, f3 p, f3 n, _Float16* act, const h4v* wl) {
  guide_encode(net, p, n, act + (threadIdx.x & 63u) * kGIn);
  guide_mlp_wave(act, wl);
}

// wave: the full forward pass; lds_x holds the encodings (64 x 48).
__device__ __forceinline__ void mlp_forward(const GuideNet& net, const _Float16* X, _Float16* H1, _Float16* H2,
                                            _Float16* H3, float* out) {
  const _Float16* w = reinterpret_cast<const _Float16*>(net.w);
  dense_fwd(X, kGIn, w + kW1, kGHid, H1, nullptr);
  dense_fwd(H1, kGHid, w + kW2, kGHid, H2, nullptr);
  dense_fwd(H2, kGHid, w + kW3, kGHid, H3, nullptr);
  dense_fwd(H3, kGHid, w + kW4, kGOut, nullptr, out);
}

// ---- mixture -----------------------------------------------------------------

__device__ __forceinline__ float vmf_norm(float k) { return k / (2.f * kGuidePi * -expm1f(-2.f * k)); }

__device__ __forceinline__ void decode_mixture(const float* raw, Mixture& m) {
  float amax = -3.0e38f;
  for (int j = 0; j < kGLobes; ++j) amax = fmaxf(amax, raw[5 * j + 4]);
  float den = 0.f;
  for (int j = 0; j < kGLobes; ++j) {
    const float* r = raw + 5 * j;
    const float nm = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + 1e-12f);
    m.mu[j] = f3{r[0] / nm, r[1] / nm, r[2] / nm};
    m.kappa[j] = expf(fminf(fmaxf(r[3], kKappaLogMin), kKappaLogMax));
    m.pi[j] = expf(r[4] - amax);
    den += m.pi[j];
  }
  for (int j = 0; j < kGLobes; ++j) m.pi[j] = m.pi[j] / den;
}

__device__ __forceinline__ float mixture_pdf(const Mixture& m, f3 w) {
  float q = 0.f;
  for (int j = 0; j < kGLobes; ++j)
    q += m.pi[j] * vmf_norm(m.kappa[j]) * expf(m.kappa[j] * (dot(m.mu[j], w) - 1.f));
  return q;
}

__device__ __forceinline__ f3 mixture_sample(const Mixture& m, float u0, float u1, float u2) {
  int j = 0;
  float cdf = m.pi[0];
  while (j < kGLobes - 1 && cdf < u0) {
    ++j;
    cdf += m.pi[j];
  }
  const float k = m.kappa[j];
  const float wz = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * k)) / k;
  const float st = sqrtf(fmaxf(0.f, 1.f - wz * wz));
  const float phi = 2.f * kGuidePi * u2;
  const Frame fr(m.mu[j]);
  return fr.world(f3{st * cosf(phi), st * sinf(phi), wz});
}

// ---- stored mixtures (one per path-queue entry) --------------------------------
//
// mix[k * C + i], k = 0..11, for queue entry i, C = WaveBuffers::mix_stride:
//   k = 0..7   {mu_j.xyz, kappa_j}
//   k = 8, 9   {pi_0..3}, {pi_4..7}
//   k = 10, 11 pi_j * C(kappa_j), the pdf weight of lobe j
// (kMixVecs = 12, guide_net.h)

// Decodes raw outputs (FP16, 5 per lobe) straight into the stored form,
// with decode_mixture's arithmetic, four values at a time.
__device__ __forceinline__ float raw_kappa(const _Float16* raw, int j) {
  return expf(fminf(fmaxf(static_cast<float>(raw[5 * j + 3]), kKappaLogMin), kKappaLogMax));
}
__device__ __forceinline__ void decode_store_mixture(const _Float16* raw, float4* mix, uint32_t C, uint32_t i) {
  float amax = -3.0e38f;
  for (int j = 0; j < kGLobes; ++j) amax = fmaxf(amax, static_cast<float>(raw[5 * j + 4]));
  float den = 0.f;
  for (int j = 0; j < kGLobes; ++j) den += expf(static_cast<float>(raw[5 * j + 4]) - amax);
  for (int j = 0; j < kGLobes; ++j) {
    const float r0 = static_cast<float>(raw[5 * j]), r1 = static_cast<float>(raw[5 * j + 1]),
                r2 = static_cast<float>(raw[5 * j + 2]);
    const float nm = sqrtf(r0 * r0 + r1 * r1 + r2 * r2 + 1e-12f);
    mix[j * C + i] = make_float4(r0 / nm, r1 / nm, r2 / nm, raw_kappa(raw, j));
  }
  for (int h = 0; h < 2; ++h) {
    float pi[4], w[4];
    for (int q = 0; q < 4; ++q) {
      const int j = 4 * h + q;
      pi[q] = expf(static_cast<float>(raw[5 * j + 4]) - amax) / den;
      w[q] = pi[q] * vmf_norm(raw_kappa(raw, j));
    }
    mix[(8 + h) * C + i] = make_float4(pi[0], pi[1], pi[2], pi[3]);
    mix[(10 + h) * C + i] = make_float4(w[0], w[1], w[2], w[3]);
  }
}

__device__ __forceinline__ float lane4(float4 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// Mixture pdf (per steradian) of the stored mixture at unit direction w.
__device__ __forceinline__ float stored_pdf(const float4* mix, uint32_t C, uint32_t i, f3 w) {
  const float4 wa = mix[10 * C + i], wb = mix[11 * C + i];
  float q = 0.f;
  for (int j = 0; j < kGLobes; ++j) {
    const float4 l = mix[j * C + i];
    q += lane4(j < 4 ? wa : wb, j & 3) * expf(l.w * (l.x * w.x + l.y * w.y + l.z * w.z - 1.f));
  }
  return q;
}

// Draws a direction from the stored mixture (lobe by CDF on u0, then
// Jakob's vMF inverse CDF on u1, u2 in the lobe's frame).
__device__ __forceinline__ f3 stored_sample(const float4* mix, uint32_t C, uint32_t i, float u0, float u1, float u2) {
  const float4 pa = mix[8 * C + i], pb = mix[9 * C + i];
  int j = 0;
  float cdf = pa.x;
  while (j < kGLobes - 1 && cdf < u0) {
    ++j;
    cdf += lane4(j < 4 ? pa : pb, j & 3);
  }
  const float4 l = mix[j * C + i];
  const float k = l.w;
  const float wz = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * k)) / k;
  const float st = sqrtf(fmaxf(0.f, 1.f - wz * wz));
  const float phi = 2.f * kGuidePi * u2;
  const Frame fr(f3{l.x, l.y, l.z});
  return fr.world(f3{st * cosf(phi), st * sinf(phi), wz});
}

// ---- training records -----------------------------------------------------------
//
// rec[k * R + r], k = 0..3, for record r of capacity R (one per guided
// vertex of a recording path):
//   {pos, T.r} {normal, T.g} {dir, T.b} {prefix.rgb, slot bits}
// T = path throughput at the vertex (after roulette); prefix = the path's
// radiance sum when its continuation reached the next vertex (written
// there), so the radiance gathered after the vertex is total - prefix.
constexpr uint32_t kNoRecord = 0xFFFFFFFFu;

// Whether the path in `slot` records training samples in batch `batch`: a
// hashed 1-in-`every` choice that changes from batch to batch.
__device__ __forceinline__ bool record_path(uint32_t slot, int batch, uint32_t every) {
  uint32_t h = slot * 0x9E3779B1u ^ (static_cast<uint32_t>(batch) * 0x85EBCA77u + 0x165667B1u);
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h % every == 0u;
}

}  // namespace xnh
