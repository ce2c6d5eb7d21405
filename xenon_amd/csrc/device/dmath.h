// dmath.h — device scalar math for the wavefront kernels.
//
// All HIP sources are compiled with -ffp-contract=off and the default
// correctly rounded fp32 division and square root, so +, -, *, /, sqrt
// round exactly as the reference's host code does (math/vec3.h:12-84,
// math/ray.h:95-113). Operators keep the reference's left-to-right
// association; that, not style, fixes their order.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace xnh {

struct f3 {
  float x, y, z;
};

// Loads through pointers known to address device (global) memory. Pointers
// read from memory (scene tables in the job list) are otherwise generic, and
// generic loads count against the LDS wait counter as well, so every LDS
// stack access would also wait for outstanding node fetches.
typedef float xv4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gld4(const float4* p, size_t i) {
  const xv4 v = ((const __attribute__((address_space(1))) xv4*)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t gld1(const uint32_t* p, size_t i) {
  return ((const __attribute__((address_space(1))) uint32_t*)p)[i];
}
__device__ __forceinline__ void gst1(uint32_t* p, size_t i, uint32_t v) {
  ((__attribute__((address_space(1))) uint32_t*)p)[i] = v;
}

// Float cos, sin, pow and exp evaluated in double and rounded once to
// float: the correctly rounded result except in rare double-rounding cases.
// The reference calls glibc's float functions (std::cos(float) etc. in
// material/bsdf.h and scene/environment.h), which compute in double
// internally and return the correctly rounded value for nearly every input;
// the device library's float versions differ from it in the last bit far
// more often, and a last-bit difference in a sampled direction can send a
// path elsewhere.
__device__ __attribute__((noinline)) float cos_r(float x) { return static_cast<float>(cos(static_cast<double>(x))); }
__device__ __attribute__((noinline)) float sin_r(float x) { return static_cast<float>(sin(static_cast<double>(x))); }
__device__ __attribute__((noinline)) float pow_r(float x, float y) {
  return static_cast<float>(pow(static_cast<double>(x), static_cast<double>(y)));
}
__device__ __attribute__((noinline)) float exp_r(float x) { return static_cast<float>(exp(static_cast<double>(x))); }

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 splat(float s) { return f3{s, s, s}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{a.x * s, a.y * s, a.z * s}; }
// Division by a scalar multiplies by its reciprocal (math/vec3.h:29).
__device__ __forceinline__ f3 operator/(f3 a, float s) {
  const float r = 1.f / s;
  return f3{a.x * r, a.y * r, a.z * r};
}
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len2(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ f3 normalize(f3 a) { return a / sqrtf(len2(a)); }
__device__ __forceinline__ f3 lerp3(f3 a, f3 b, float t) { return a + t * (b - a); }
__device__ __forceinline__ f3 reflect(f3 v, f3 n) { return v - 2.f * dot(v, n) * n; }
__device__ __forceinline__ float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// std::max / std::min / std::clamp value semantics (first operand kept on
// ties and NaN comparisons), used wherever the reference calls them.
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float sclamp(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
// SSE MINPS / MAXPS lane semantics (geometry/blas.cpp:40-70).
__device__ __forceinline__ float xmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float xmax(float a, float b) { return a > b ? a : b; }

__device__ __forceinline__ bool finite3(f3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

// PCG32 XSH-RR (camera/sampler.h:13-36), stream increment 1.
struct Rng {
  uint64_t s;
  __device__ __forceinline__ uint32_t u32() {
    const uint64_t old = s;
    s = old * 6364136223846793005ULL + 1ULL;
    const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
  }
  __device__ __forceinline__ float f() { return static_cast<float>(u32() >> 8) * (1.f / 16777216.f); }
};

// seed_pcg(seed, 0): state 0, one step, add the seed, one step.
__device__ __forceinline__ Rng rng_for(uint64_t seed) {
  Rng r{0};
  r.u32();
  r.s += seed;
  r.u32();
  return r;
}

// Duff et al. 2017 frame (math/ray.h:95-113).
struct Frame {
  f3 t, b, n;
  __device__ __forceinline__ explicit Frame(f3 nn) : n(nn) {
    const float s = (n.z >= 0.f) ? 1.f : -1.f;
    const float a = -1.f / (s + n.z);
    const float c = n.x * n.y * a;
    t = f3{1.f + s * n.x * n.x * a, s * c, -s * n.x};
    b = f3{c, s + n.y * n.y * a, -n.y};
  }
  __device__ __forceinline__ f3 world(f3 l) const { return l.x * t + l.y * b + l.z * n; }
  __device__ __forceinline__ f3 local(f3 w) const { return f3{dot(w, t), dot(w, b), dot(w, n)}; }
};

// Offset along the geometric normal toward `dir` (math/ray.h:22-26).
__device__ __forceinline__ f3 offset_origin(f3 p, f3 ng, f3 dir, float eps) {
  f3 n = ng;
  if (dot(n, dir) < 0.f) n = -n;
  return p + n * eps;
}

__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }

}  // namespace xnh
