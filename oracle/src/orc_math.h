// orc_math.h — scalar math core of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY. The oracle is a plain-C++ restatement of the
// reference CPU path tracer; it is used by tests/ and by bench.py's CPU
// baseline leg as the checker, never by the product library.
//
// Every operation below keeps the reference's IEEE-754 single-precision
// evaluation order (left-to-right, no contraction: the oracle is compiled
// with -ffp-contract=off), because the GPU path is held to bit-exact hit
// parity against it.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

namespace orc {

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr float kPi = 3.14159265358979323846f;      // math/ray.h:15
constexpr float kInvPi = 1.f / kPi;                  // math/ray.h:16
constexpr float kBounceEps = 1e-4f;                  // math/ray.h:13
constexpr float kShadowOffset = 5e-4f;               // math/ray.h:14

// Three-float vector. Operators follow math/vec3.h:12-84: component-wise,
// division by a scalar multiplies by its reciprocal (vec3.h:29).
struct F3 {
  float x = 0.f, y = 0.f, z = 0.f;
  F3() = default;
  constexpr F3(float a, float b, float c) : x(a), y(b), z(c) {}
  static F3 splat(float s) { return F3(s, s, s); }
  float at(int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline F3 operator+(F3 a, F3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator*(F3 a, F3 b) { return F3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline F3 operator/(F3 a, F3 b) { return F3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline F3 operator*(F3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
inline F3 operator*(float s, F3 a) { return a * s; }
inline F3 operator/(F3 a, float s) {
  const float r = 1.f / s;
  return F3(a.x * r, a.y * r, a.z * r);
}
inline F3 operator-(F3 a) { return F3(-a.x, -a.y, -a.z); }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) {
  return F3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float len_sq(F3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float length(F3 a) { return std::sqrt(len_sq(a)); }
inline F3 normalize(F3 a) { return a / length(a); }
inline F3 lerp(F3 a, F3 b, float t) { return a + t * (b - a); }   // vec3.h:54
inline F3 reflect(F3 v, F3 n) { return v - 2.f * dot(v, n) * n; } // vec3.h:52
// First-greatest / first-smallest folds, the semantics of
// std::max({...}) and std::min({...}) (libstdc++ max_element/min_element).
inline float max_first(float a, float b) { return (a < b) ? b : a; }
inline float min_first(float a, float b) { return (b < a) ? b : a; }
inline float max_comp(F3 v) { return max_first(max_first(v.x, v.y), v.z); } // vec3.h:72
inline int max_axis(F3 v) { return (v.x > v.y && v.x > v.z) ? 0 : (v.y > v.z ? 1 : 2); } // vec3.h:73-75
// SSE MINPS / MAXPS lane semantics: the second operand wins unless the
// comparison holds (so a NaN in either operand yields the second one).
inline float sse_min(float a, float b) { return a < b ? a : b; }
inline float sse_max(float a, float b) { return a > b ? a : b; }

// Snell refraction, vec3.h:61-68.
inline bool refract_dir(F3 v, F3 n, float eta, F3& out) {
  const float ci = -dot(normalize(v), n);
  const float s2 = eta * eta * (1.f - ci * ci);
  if (s2 >= 1.f) return false;
  out = eta * v + (eta * ci - std::sqrt(1.f - s2)) * n;
  return true;
}

// PCG32 XSH-RR with the reference's seeding (camera/sampler.h:13-36).
// The stream increment is (seq<<1)|1; the integrator always uses seq 0, so
// the increment is 1.
struct Pcg {
  uint64_t state = 0;
  uint64_t inc = 1;
  uint32_t next_u32() {
    const uint64_t old = state;
    state = old * 6364136223846793005ULL + (inc | 1ULL);
    const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = static_cast<uint32_t>(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
  }
  float next_f32() { return static_cast<float>(next_u32() >> 8) * (1.f / 16777216.f); }
};
inline Pcg pcg_seeded(uint64_t seed, uint64_t seq = 0) {
  Pcg p;
  p.state = 0;
  p.inc = (seq << 1u) | 1u;
  p.next_u32();
  p.state += seed;
  p.next_u32();
  return p;
}

// Frisvad/Duff orthonormal frame (math/ray.h:95-113).
struct Frame {
  F3 t, b, n;
  explicit Frame(F3 nrm) : n(nrm) {
    const float s = (n.z >= 0.f) ? 1.f : -1.f;
    const float a = -1.f / (s + n.z);
    const float c = n.x * n.y * a;
    t = F3(1.f + s * n.x * n.x * a, s * c, -s * n.x);
    b = F3(c, s + n.y * n.y * a, -n.y);
  }
  F3 to_world(F3 l) const { return l.x * t + l.y * b + l.z * n; }
  F3 to_local(F3 w) const { return F3(dot(w, t), dot(w, b), dot(w, n)); }
};

// Nudge a surface point along the geometric normal toward `dir`
// (math/ray.h:22-26).
inline F3 nudge(F3 p, F3 ng, F3 dir, float eps) {
  F3 n = ng;
  if (dot(n, dir) < 0.f) n = -n;
  return p + n * eps;
}

struct RayF {
  F3 o, d;
  float tmin = kBounceEps;
  float tmax = kInf;
  F3 at(float t) const { return o + t * d; }
};

struct HitF {
  float t = kInf;
  F3 pos, n, gn;
  int mat = -1, prim = -1, inst = -1;
  float u = 0.f, v = 0.f;
  bool front = true;
};

}  // namespace orc
