// dbsdf.h — principled BSDF on the device: diffuse, GGX conductor, GGX and
// delta dielectric, the diffuse-like subsurface lobe, NEE eval/pdf and lobe
// selection, with the semantics of material/bsdf.h:30-708.
//
// RNG draw order: the reference passes two draws straight into a call
// (bsdf.h:252, 347, 429, 566) and is built with g++, which evaluates those
// call arguments right to left, so the SECOND draw becomes u1. The kernels
// draw explicitly in that order.
//
// Transcendentals (cos, sin, pow, exp) are the device's and may differ from
// glibc by an ulp; everything else rounds as the host does.
#pragma once
#include "../common/dev_types.h"
#include "dmath.h"
#include "xenon_hip.h"

namespace xnh {

// Material as six float4 in HBM (kMatVecs):
//   {base.rgb, metallic} {ior, alpha, alpha_x, alpha_y} {f0.rgb, subsurface}
//   {subsurface.rgb, flags} {ms_k1, ms_k2, ms_f.r, ms_f.g} {ms_f.b, 0, 0, 0}
// The last two hold the parts of the multiple-scattering compensation
// (bsdf.h:212-249) that depend on the material alone, computed at upload
// with the same float operations in the same order: ms_k1 = 5 exp(-2.69 a),
// ms_k2 = 1 + 22.7 a^1.5 and ms_f = favg^2 E_avg / max(1 - favg.r (1 - E_avg), 1e-4).
struct DMat {
  f3 base, f0, ss, ms_f;
  float metallic, ior, alpha, ax, ay, sss, ms_k1, ms_k2;
  uint32_t flags;
  __device__ __forceinline__ bool conductor() const { return flags & XNH_MAT_CONDUCTOR; }
  __device__ __forceinline__ bool transmissive() const { return flags & XNH_MAT_TRANSMISSIVE; }
  __device__ __forceinline__ bool has_sss() const { return flags & XNH_MAT_SUBSURFACE; }
  __device__ __forceinline__ bool delta() const { return flags & XNH_MAT_DELTA; }
};

__device__ __forceinline__ DMat load_mat(const float4* mats, int id) {
  const float4* r = mats + kMatVecs * id;
  const float4 a = r[0], b = r[1], c = r[2], d = r[3], e = r[4], g = r[5];
  DMat m;
  m.base = xyz(a);
  m.metallic = a.w;
  m.ior = b.x;
  m.alpha = b.y;
  m.ax = b.z;
  m.ay = b.w;
  m.f0 = xyz(c);
  m.sss = c.w;
  m.ss = xyz(d);
  m.flags = __float_as_uint(d.w);
  m.ms_k1 = e.x;
  m.ms_k2 = e.y;
  m.ms_f = f3{e.z, e.w, g.x};
  return m;
}

struct BSample {
  f3 wi, f;
  float pdf;
  bool delta;
};

__device__ __forceinline__ float c2t(f3 w) { return w.z * w.z; }
__device__ __forceinline__ float s2t(f3 w) { return smax(0.f, 1.f - c2t(w)); }
__device__ __forceinline__ float st_(f3 w) { return sqrtf(s2t(w)); }
__device__ __forceinline__ float t2t(f3 w) {
  const float c2 = c2t(w);
  return (c2 < 1e-8f) ? 0.f : s2t(w) / c2;
}
__device__ __forceinline__ float cphi(f3 w) {
  const float s = st_(w);
  return (s == 0.f) ? 1.f : sclamp(w.x / s, -1.f, 1.f);
}
__device__ __forceinline__ float sphi(f3 w) {
  const float s = st_(w);
  return (s == 0.f) ? 0.f : sclamp(w.y / s, -1.f, 1.f);
}

__device__ __forceinline__ float mis_power2(float a, float b) {
  const float a2 = a * a, b2 = b * b, s = a2 + b2;
  return (s > 0.f) ? a2 / s : 0.f;
}

constexpr float kPiD = 3.14159265358979323846f;
constexpr float kInvPiD = 1.f / kPiD;

__device__ __forceinline__ float D_ggx(f3 h, float ax, float ay) {
  const float c2 = c2t(h);
  if (c2 <= 0.f) return 0.f;
  const float ex = cphi(h) / ax, ey = sphi(h) / ay;
  const float tn = (ex * ex + ey * ey) * t2t(h);
  const float d = c2 * c2 * (1.f + tn);
  return 1.f / (kPiD * ax * ay * d * d);
}
__device__ __forceinline__ float Lambda_ggx(f3 w, float ax, float ay) {
  const float t2 = t2t(w);
  if (t2 == 0.f) return 0.f;
  const float cp = cphi(w), sp = sphi(w);
  const float a2 = (cp * ax) * (cp * ax) + (sp * ay) * (sp * ay);
  return (-1.f + sqrtf(1.f + a2 * t2)) * 0.5f;
}
__device__ __forceinline__ float G1_ggx(f3 w, float ax, float ay) { return 1.f / (1.f + Lambda_ggx(w, ax, ay)); }
__device__ __forceinline__ float G2_ggx(f3 o, f3 i, float ax, float ay) {
  return 1.f / (1.f + Lambda_ggx(o, ax, ay) + Lambda_ggx(i, ax, ay));
}

__device__ __forceinline__ float fresnel_d(float ci, float ior) {
  ci = sclamp(ci, -1.f, 1.f);
  float ei = 1.f, et = ior;
  if (ci < 0.f) {
    const float tmp = ei;
    ei = et;
    et = tmp;
    ci = -ci;
  }
  const float s2i = 1.f - ci * ci;
  const float s2o = (ei / et) * (ei / et) * s2i;
  if (s2o >= 1.f) return 1.f;
  const float ct = sqrtf(smax(0.f, 1.f - s2o));
  const float rpa = ((et * ci) - (ei * ct)) / ((et * ci) + (ei * ct));
  const float rpe = ((ei * ci) - (et * ct)) / ((ei * ci) + (et * ct));
  return (rpa * rpa + rpe * rpe) * 0.5f;
}
__device__ __forceinline__ f3 fresnel_s(float c, f3 f0) {
  const float m = sclamp(1.f - c, 0.f, 1.f);
  const float m2 = m * m;
  return f0 + (splat(1.f) - f0) * (m2 * m2 * m);
}

__device__ __forceinline__ f3 cosine_dir(float u1, float u2) {
  const float r = sqrtf(u1);
  const float phi = 2.0f * kPiD * u2;
  return f3{r * cos_r(phi), r * sin_r(phi), sqrtf(smax(0.0f, 1.0f - u1))};
}

// Heitz 2018 visible-normal sampling (bsdf.h:164-190).
__device__ __forceinline__ f3 vndf(f3 wo, float ax, float ay, float u1, float u2) {
  const f3 vh = normalize(f3{ax * wo.x, ay * wo.y, wo.z});
  const float l2 = vh.x * vh.x + vh.y * vh.y;
  const f3 t1 = (l2 > 1e-7f) ? f3{-vh.y, vh.x, 0.f} * (1.f / sqrtf(l2)) : f3{1.f, 0.f, 0.f};
  const f3 t2 = cross(vh, t1);
  const float r = sqrtf(u1);
  const float phi = 2.0f * kPiD * u2;
  const float p1 = r * cos_r(phi);
  float p2 = r * sin_r(phi);
  const float s = 0.5f * (1.0f + vh.z);
  p2 = (1.0f - s) * sqrtf(smax(0.f, 1.0f - p1 * p1)) + s * p2;
  const f3 nh = p1 * t1 + p2 * t2 + sqrtf(smax(0.f, 1.0f - p1 * p1 - p2 * p2)) * vh;
  return normalize(f3{ax * nh.x, ay * nh.y, smax(0.f, nh.z)});
}
__device__ __forceinline__ float vndf_pdf(f3 wo, f3 h, float ax, float ay) {
  const float D = D_ggx(h, ax, ay);
  const float g = G1_ggx(wo, ax, ay);
  const float co = fabsf(wo.z);
  if (co < 1e-8f) return 0.f;
  return D * g * smax(0.f, dot(wo, h)) / co;
}

// bsdf.h:212-249 per evaluation: the directional albedo E(c) and the
// compensation factor, with the material's constant parts precomputed.
__device__ __forceinline__ float ms_E(float c, const DMat& m) {
  const float e = 1.0f - pow_r(1.0f - c, m.ms_k1) / m.ms_k2;
  return sclamp(e, 0.f, 1.f);
}
__device__ __forceinline__ f3 ms_comp(const DMat& m, float co, float ci) {
  const float eo = ms_E(fabsf(co), m), ei = ms_E(fabsf(ci), m);
  const float c = (1.f - eo) * (1.f - ei);
  return splat(1.f) + m.ms_f * c / splat(smax(eo * ei, 1e-4f));
}

// ---- lobes ------------------------------------------------------------------
__device__ __forceinline__ f3 diffuse_f(f3 wi, const DMat& m) {
  if (wi.z <= 0.f) return splat(0.f);
  return m.base * (1.f - m.metallic) * kInvPiD;
}
__device__ __forceinline__ float diffuse_pdf(f3 wi) { return (wi.z <= 0.f) ? 0.f : wi.z * kInvPiD; }

__device__ __forceinline__ f3 refl_F(float ch, const DMat& m) {
  if (m.conductor()) return fresnel_s(fabsf(ch), m.f0);
  return splat(fresnel_d(ch, m.ior));
}
__device__ __forceinline__ f3 refl_f(f3 wo, f3 wi, const DMat& m) {
  if (wo.z <= 0.f || wi.z <= 0.f) return splat(0.f);
  const f3 h = normalize(wo + wi);
  if (h.z <= 0.f) return splat(0.f);
  const float D = D_ggx(h, m.ax, m.ay);
  const float G = G2_ggx(wo, wi, m.ax, m.ay);
  const f3 F = refl_F(dot(wo, h), m);
  const float den = 4.f * wo.z * wi.z;
  if (den < 1e-8f) return splat(0.f);
  f3 r = F * D * G / den;
  if (!m.transmissive()) r = r * ms_comp(m, wo.z, wi.z);
  return r;
}
__device__ __forceinline__ float refl_pdf(f3 wo, f3 wi, const DMat& m) {
  if (wo.z <= 0.f || wi.z <= 0.f) return 0.f;
  const f3 h = normalize(wo + wi);
  if (h.z <= 0.f) return 0.f;
  return vndf_pdf(wo, h, m.ax, m.ay) * (1.f / (4.f * fabsf(dot(wo, h))));
}
__device__ __forceinline__ f3 trans_f(f3 wo, f3 wi, const DMat& m) {
  if (wo.z == 0.f || wi.z == 0.f) return splat(0.f);
  if (wo.z * wi.z > 0.f) return splat(0.f);
  const float eta = (wo.z > 0.f) ? (1.f / m.ior) : m.ior;
  f3 h = normalize(wo + eta * wi);
  if (h.z < 0.f) h = -h;
  const float oh = dot(wo, h), ih = dot(wi, h);
  if (oh * ih >= 0.f) return splat(0.f);
  const float D = D_ggx(h, m.ax, m.ay);
  const float G = G2_ggx(wo, wi, m.ax, m.ay);
  const float F = fresnel_d(oh, m.ior);
  const float co = fabsf(wo.z), ci = fabsf(wi.z);
  if (co < 1e-8f || ci < 1e-8f) return splat(0.f);
  const float den = oh + eta * ih;
  const float den2 = den * den;
  if (fabsf(den) < 1e-8f) return splat(0.f);
  const float fac = fabsf(oh * ih) * (eta * eta) / (co * ci * den2);
  return m.base * ((1.f - F) * D * G * fac);
}
__device__ __forceinline__ float trans_pdf(f3 wo, f3 wi, const DMat& m) {
  if (wo.z == 0.f || wi.z == 0.f) return 0.f;
  if (wo.z * wi.z > 0.f) return 0.f;
  const float eta = (wo.z > 0.f) ? (1.f / m.ior) : m.ior;
  f3 h = normalize(wo + eta * wi);
  if (h.z < 0.f) h = -h;
  const float oh = dot(wo, h), ih = dot(wi, h);
  if (oh * ih >= 0.f) return 0.f;
  const float ph = vndf_pdf(wo, h, m.ax, m.ay);
  const float den = oh + eta * ih;
  if (fabsf(den) < 1e-8f) return 0.f;
  return ph * fabsf((eta * eta * ih) / (den * den));
}
__device__ __forceinline__ f3 sss_f(f3 wi, const DMat& m) {
  if (wi.z <= 0.f) return splat(0.f);
  return lerp3(m.base, m.ss, m.sss) * (1.f - m.metallic) * kInvPiD;
}

// True when the BSDF is exactly Lambertian: an opaque, non-metal dielectric
// with F0 = 0 (IOR 1), whose sampler always takes the cosine lobe
// (bsdf.h:666-708 with a zero specular weight) and whose NEE terms reduce to
// the diffuse lobe. Path guiding mixes its sampler with this one.
__device__ __forceinline__ bool lambert_exact(const DMat& m) {
  return (m.flags & (XNH_MAT_DELTA | XNH_MAT_CONDUCTOR | XNH_MAT_TRANSMISSIVE | XNH_MAT_SUBSURFACE)) == 0u &&
         m.f0.x == 0.f;
}

// NEE evaluation summed over non-delta lobes (bsdf.h:601-657).
__device__ __forceinline__ f3 nee_f(f3 wo, f3 wi, const DMat& m) {
  if (m.delta()) return splat(0.f);
  if (m.conductor()) return refl_f(wo, wi, m);
  if (m.transmissive()) return (wi.z > 0.f) ? refl_f(wo, wi, m) : trans_f(wo, wi, m);
  if (m.has_sss()) return sss_f(wi, m);
  const float ws = m.f0.x, wd = 1.f - ws;
  if (wi.z > 0.f) return diffuse_f(wi, m) * wd + refl_f(wo, wi, m) * ws;
  return splat(0.f);
}
__device__ __forceinline__ float nee_pdf(f3 wo, f3 wi, const DMat& m) {
  if (m.delta()) return 0.f;
  if (m.conductor()) return refl_pdf(wo, wi, m);
  if (m.transmissive()) {
    const float F = fresnel_d(fabsf(wo.z), m.ior);
    if (wi.z > 0.f) return F * refl_pdf(wo, wi, m);
    return (1.f - F) * trans_pdf(wo, wi, m);
  }
  if (m.has_sss()) return diffuse_pdf(wi);
  const float ws = m.f0.x, wd = 1.f - ws;
  if (wi.z > 0.f) return diffuse_pdf(wi) * wd + refl_pdf(wo, wi, m) * ws;
  return 0.f;
}

__device__ __forceinline__ bool sample_diffuse_like(f3 wo, f3 color, const DMat& m, Rng& rng, BSample& s) {
  const float first = rng.f();
  const float second = rng.f();
  s.wi = cosine_dir(second, first);
  if (wo.z < 0.f) s.wi.z = -s.wi.z;
  s.pdf = fabsf(s.wi.z) * kInvPiD;
  if (s.pdf < 1e-8f) return false;
  s.f = color * (1.f - m.metallic) * kInvPiD;
  s.delta = false;
  return true;
}
__device__ __forceinline__ bool sample_refl(f3 wo, const DMat& m, Rng& rng, BSample& s) {
  if (wo.z <= 0.f) return false;
  const float first = rng.f();
  const float second = rng.f();
  const f3 h = vndf(wo, m.ax, m.ay, second, first);
  if (h.z <= 0.f) return false;
  s.wi = reflect(-wo, h);
  if (s.wi.z <= 0.f) return false;
  s.pdf = refl_pdf(wo, s.wi, m);
  if (s.pdf < 1e-8f) return false;
  s.f = refl_f(wo, s.wi, m);
  s.delta = false;
  return true;
}
__device__ __forceinline__ bool sample_trans(f3 wo, const DMat& m, Rng& rng, BSample& s) {
  if (fabsf(wo.z) < 1e-8f) return false;
  const float first = rng.f();
  const float second = rng.f();
  f3 h = vndf(wo, m.ax, m.ay, second, first);
  if (h.z < 0.f) h = -h;
  const float eta = (wo.z > 0.f) ? (1.f / m.ior) : m.ior;
  const float oh = dot(wo, h);
  const float s2o = smax(0.f, 1.f - oh * oh);
  const float s2i = eta * eta * s2o;
  if (s2i >= 1.f) return false;
  const float ci = sqrtf(smax(0.f, 1.f - s2i));
  const float sg = (oh > 0.f) ? 1.f : -1.f;
  s.wi = normalize((-eta * wo) + (eta * oh - sg * ci) * h);
  if (wo.z * s.wi.z > 0.f) return false;
  s.pdf = trans_pdf(wo, s.wi, m);
  if (s.pdf < 1e-8f) return false;
  s.f = trans_f(wo, s.wi, m);
  s.delta = false;
  return true;
}
__device__ __forceinline__ bool sample_mirror(f3 wo, const DMat& m, BSample& s) {
  if (fabsf(wo.z) < 1e-8f) return false;
  const f3 n{0.f, 0.f, (wo.z > 0.f) ? 1.f : -1.f};
  s.wi = reflect(-wo, n);
  const f3 F = m.conductor() ? fresnel_s(fabsf(wo.z), m.f0) : splat(fresnel_d(fabsf(wo.z), m.ior));
  const float ci = fabsf(s.wi.z);
  if (ci < 1e-8f) return false;
  s.f = F / ci;
  s.pdf = 1.f;
  s.delta = true;
  return true;
}
__device__ __forceinline__ bool sample_glass(f3 wo, const DMat& m, BSample& s) {
  if (fabsf(wo.z) < 1e-8f) return false;
  const float eta = (wo.z > 0.f) ? (1.f / m.ior) : m.ior;
  const f3 n{0.f, 0.f, (wo.z > 0.f) ? 1.f : -1.f};
  // Snell refraction of v = -wo about n (math/vec3.h:61-68)
  const f3 v = -wo;
  const float ci0 = -dot(normalize(v), n);
  const float s2 = eta * eta * (1.f - ci0 * ci0);
  if (s2 >= 1.f) return false;
  s.wi = normalize(eta * v + (eta * ci0 - sqrtf(1.f - s2)) * n);
  const float F = fresnel_d(fabsf(wo.z), m.ior);
  const float ci = fabsf(s.wi.z);
  if (ci < 1e-8f) return false;
  s.f = m.base * ((1.f - F) * eta * eta / ci);
  s.pdf = 1.f;
  s.delta = true;
  return true;
}

// Lobe selection (bsdf.h:666-708).
__device__ __forceinline__ bool sample_bsdf(f3 wo, const DMat& m, Rng& rng, BSample& s) {
  if (m.delta()) {
    if (m.transmissive()) {
      const float F = fresnel_d(wo.z, m.ior);
      if (rng.f() < F) return sample_mirror(wo, m, s);
      return sample_glass(wo, m, s);
    }
    return sample_mirror(wo, m, s);
  }
  if (m.conductor()) return sample_refl(wo, m, rng, s);
  if (m.has_sss()) return sample_diffuse_like(wo, lerp3(m.base, m.ss, m.sss), m, rng, s);
  if (m.transmissive()) {
    const float F = fresnel_d(fabsf(wo.z), m.ior);
    if (rng.f() < F) {
      if (!sample_refl(wo, m, rng, s)) return false;
      s.pdf *= F;
      return true;
    }
    if (!sample_trans(wo, m, rng, s)) return false;
    s.pdf *= (1.f - F);
    return true;
  }
  const float ws = sclamp(m.f0.x, 0.f, 1.f);
  const float wd = 1.f - ws;
  if (rng.f() < ws) {
    if (!sample_refl(wo, m, rng, s)) return false;
    s.pdf *= ws;
    return true;
  }
  if (!sample_diffuse_like(wo, m.base, m, rng, s)) return false;
  s.pdf *= wd;
  return true;
}

}  // namespace xnh
