// orc_bsdf.h — material preparation and principled BSDF of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
//
// Restates material/material.h:55-75 and material/bsdf.h:30-708 in the
// reference's float evaluation order. Subsurface is kept because the
// reference's lobe dispatch consults it (bsdf.h:683-684).
//
// RNG draw order: the reference passes two draws straight into a call,
// `f(rng.next_float(), rng.next_float())` (bsdf.h:252, 347, 429, 566). The
// reference is built with g++, which evaluates those arguments right to
// left, so the SECOND draw becomes the first parameter (u1). The oracle
// draws explicitly in that order.
#pragma once
#include <cmath>

#include "orc_math.h"

namespace orc {

inline float stdmax(float a, float b) { return (a < b) ? b : a; }
inline float stdmin(float a, float b) { return (b < a) ? b : a; }
inline float stdclamp(float v, float lo, float hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

struct Mat {
  // authoring parameters (material.h:24-37)
  F3 base = F3::splat(0.5f);
  float roughness = 0.5f, metallic = 0.f, ior = 1.5f, transmission = 0.f, subsurface = 0.f;
  F3 ss_color = F3::splat(1.f);
  float anisotropy = 0.f;
  F3 emission = F3::splat(0.f);
  // derived (material.h:40-48)
  float alpha = 0.25f, ax = 0.25f, ay = 0.25f;
  F3 f0 = F3::splat(0.04f);
  bool conductor = false, transmissive = false, sss = false, delta = false, emissive = false;
};

// material.h:55-75
inline void prepare(Mat& m) {
  m.alpha = stdmax(m.roughness * m.roughness, 1e-6f);
  const float aspect = std::sqrt(1.0f - 0.9f * std::fabs(m.anisotropy));
  m.ax = stdmax(m.alpha / aspect, 1e-6f);
  m.ay = stdmax(m.alpha * aspect, 1e-6f);
  const float fd = ((m.ior - 1.0f) * (m.ior - 1.0f)) / ((m.ior + 1.0f) * (m.ior + 1.0f));
  m.f0 = lerp(F3::splat(fd), m.base, m.metallic);
  m.conductor = m.metallic > 0.5f;
  m.transmissive = m.transmission > 0.01f;
  m.sss = m.subsurface > 0.01f;
  m.delta = m.alpha < 0.001f;
  m.emissive = m.emission.x > 0.0f || m.emission.y > 0.0f || m.emission.z > 0.0f;
}

struct Sample {
  F3 wi, f;
  float pdf = 0.f;
  bool delta = false;
};

// ---- local-frame trigonometry (bsdf.h:30-43) ------------------------------
inline float cos2t(F3 w) { return w.z * w.z; }
inline float sin2t(F3 w) { return stdmax(0.f, 1.f - cos2t(w)); }
inline float sint(F3 w) { return std::sqrt(sin2t(w)); }
inline float tan2t(F3 w) {
  const float c2 = cos2t(w);
  return (c2 < 1e-8f) ? 0.f : sin2t(w) / c2;
}
inline float cosp(F3 w) {
  const float s = sint(w);
  return (s == 0.f) ? 1.f : stdclamp(w.x / s, -1.f, 1.f);
}
inline float sinp(F3 w) {
  const float s = sint(w);
  return (s == 0.f) ? 0.f : stdclamp(w.y / s, -1.f, 1.f);
}

// power heuristic, bsdf.h:46-51
inline float power2(float a, float b) {
  const float a2 = a * a, b2 = b * b, s = a2 + b2;
  return (s > 0.f) ? a2 / s : 0.f;
}

// anisotropic GGX D (bsdf.h:68-76)
inline float ggx_D(F3 h, float ax, float ay) {
  const float c2 = cos2t(h);
  if (c2 <= 0.f) return 0.f;
  const float ex = cosp(h) / ax, ey = sinp(h) / ay;
  const float t2 = (ex * ex + ey * ey) * tan2t(h);
  const float d = c2 * c2 * (1.f + t2);
  return 1.f / (kPi * ax * ay * d * d);
}
// anisotropic Smith Lambda (bsdf.h:91-97) and height-correlated G2 (bsdf.h:107-109)
inline float ggx_Lambda(F3 w, float ax, float ay) {
  const float t2 = tan2t(w);
  if (t2 == 0.f) return 0.f;
  const float cp = cosp(w), sp = sinp(w);
  const float a2 = (cp * ax) * (cp * ax) + (sp * ay) * (sp * ay);
  return (-1.f + std::sqrt(1.f + a2 * t2)) * 0.5f;
}
inline float ggx_G1(F3 w, float ax, float ay) { return 1.f / (1.f + ggx_Lambda(w, ax, ay)); }
inline float ggx_G2(F3 o, F3 i, float ax, float ay) {
  return 1.f / (1.f + ggx_Lambda(o, ax, ay) + ggx_Lambda(i, ax, ay));
}

// exact dielectric Fresnel with TIR (bsdf.h:115-131)
inline float fr_dielectric(float ci, float ior) {
  ci = stdclamp(ci, -1.f, 1.f);
  float ei = 1.f, et = ior;
  if (ci < 0.f) {
    const float tmp = ei; ei = et; et = tmp;
    ci = -ci;
  }
  const float s2i = 1.f - ci * ci;
  const float s2t = (ei / et) * (ei / et) * s2i;
  if (s2t >= 1.f) return 1.f;
  const float ct = std::sqrt(stdmax(0.f, 1.f - s2t));
  const float rpa = ((et * ci) - (ei * ct)) / ((et * ci) + (ei * ct));
  const float rpe = ((ei * ci) - (et * ct)) / ((ei * ci) + (et * ct));
  return (rpa * rpa + rpe * rpe) * 0.5f;
}
// Schlick with RGB F0 (bsdf.h:134-138)
inline F3 fr_schlick(float c, F3 f0) {
  const float m = stdclamp(1.f - c, 0.f, 1.f);
  const float m2 = m * m;
  return f0 + (F3::splat(1.f) - f0) * (m2 * m2 * m);
}

// cosine hemisphere (bsdf.h:151-156)
inline F3 cos_hemi(float u1, float u2) {
  const float r = std::sqrt(u1);
  const float phi = 2.0f * kPi * u2;
  return F3(r * std::cos(phi), r * std::sin(phi), std::sqrt(stdmax(0.0f, 1.0f - u1)));
}

// visible-normal sampling, Heitz 2018 (bsdf.h:164-190)
inline F3 vndf_sample(F3 wo, float ax, float ay, float u1, float u2) {
  const F3 vh = normalize(F3(ax * wo.x, ay * wo.y, wo.z));
  const float l2 = vh.x * vh.x + vh.y * vh.y;
  const F3 t1v = (l2 > 1e-7f) ? F3(-vh.y, vh.x, 0.f) * (1.f / std::sqrt(l2)) : F3(1.f, 0.f, 0.f);
  const F3 t2v = cross(vh, t1v);
  const float r = std::sqrt(u1);
  const float phi = 2.0f * kPi * u2;
  const float p1 = r * std::cos(phi);
  float p2 = r * std::sin(phi);
  const float s = 0.5f * (1.0f + vh.z);
  p2 = (1.0f - s) * std::sqrt(stdmax(0.f, 1.0f - p1 * p1)) + s * p2;
  const F3 nh = p1 * t1v + p2 * t2v + std::sqrt(stdmax(0.f, 1.0f - p1 * p1 - p2 * p2)) * vh;
  return normalize(F3(ax * nh.x, ay * nh.y, stdmax(0.f, nh.z)));
}
inline float vndf_pdf(F3 wo, F3 h, float ax, float ay) {
  const float D = ggx_D(h, ax, ay);
  const float g1 = ggx_G1(wo, ax, ay);
  const float co = std::fabs(wo.z);
  if (co < 1e-8f) return 0.f;
  return D * g1 * stdmax(0.f, dot(wo, h)) / co;
}

// multiple-scattering energy fit (bsdf.h:212-249)
inline float ms_albedo(float c, float a) {
  const float e = 1.0f - std::pow(1.0f - c, 5.0f * std::exp(-2.69f * a)) / (1.0f + 22.7f * std::pow(a, 1.5f));
  return stdclamp(e, 0.f, 1.f);
}
inline float ms_avg(float a) { return 1.0f / (1.0f + 0.3037f * a + 0.1218f * a * a); }
inline F3 ms_factor(F3 f0, float co, float ci, float a) {
  const float eo = ms_albedo(std::fabs(co), a);
  const float ei = ms_albedo(std::fabs(ci), a);
  const float ea = ms_avg(a);
  const F3 favg = (F3::splat(1.f) + f0 * 20.f) * (1.f / 21.f);
  float den = 1.f - favg.x * (1.f - ea);
  den = stdmax(den, 1e-4f);
  const F3 fms = favg * favg * ea / F3::splat(den);
  const float comp = (1.f - eo) * (1.f - ei);
  return F3::splat(1.f) + fms * comp / F3::splat(stdmax(eo * ei, 1e-4f));
}

// ---- lobes ------------------------------------------------------------------
// Lambert (bsdf.h:256-279)
inline F3 lambert_f(F3 wi, const Mat& m) {
  if (wi.z <= 0.f) return F3::splat(0.f);
  return m.base * (1.f - m.metallic) * kInvPi;
}
inline float lambert_pdf(F3 wi) { return (wi.z <= 0.f) ? 0.f : wi.z * kInvPi; }
inline bool lambert_sample(F3 wo, const Mat& m, Pcg& rng, Sample& s) {
  const float first = rng.next_f32();
  const float second = rng.next_f32();
  s.wi = cos_hemi(second, first);
  if (wo.z < 0.f) s.wi.z = -s.wi.z;
  s.pdf = std::fabs(s.wi.z) * kInvPi;
  if (s.pdf < 1e-8f) return false;
  s.f = m.base * (1.f - m.metallic) * kInvPi;
  s.delta = false;
  return true;
}

// GGX reflection (bsdf.h:288-356)
inline F3 ggxr_fresnel(float ch, const Mat& m) {
  if (m.conductor) return fr_schlick(std::fabs(ch), m.f0);
  return F3::splat(fr_dielectric(ch, m.ior));
}
inline F3 ggxr_f(F3 wo, F3 wi, const Mat& m) {
  if (wo.z <= 0.f || wi.z <= 0.f) return F3::splat(0.f);
  const F3 h = normalize(wo + wi);
  if (h.z <= 0.f) return F3::splat(0.f);
  const float D = ggx_D(h, m.ax, m.ay);
  const float G = ggx_G2(wo, wi, m.ax, m.ay);
  const F3 F = ggxr_fresnel(dot(wo, h), m);
  const float den = 4.f * wo.z * wi.z;
  if (den < 1e-8f) return F3::splat(0.f);
  F3 r = F * D * G / den;
  if (!m.transmissive) r = r * ms_factor(m.f0, wo.z, wi.z, m.alpha);
  return r;
}
inline float ggxr_pdf(F3 wo, F3 wi, const Mat& m) {
  if (wo.z <= 0.f || wi.z <= 0.f) return 0.f;
  const F3 h = normalize(wo + wi);
  if (h.z <= 0.f) return 0.f;
  const float dv = vndf_pdf(wo, h, m.ax, m.ay);
  const float jac = 1.f / (4.f * std::fabs(dot(wo, h)));
  return dv * jac;
}
inline bool ggxr_sample(F3 wo, const Mat& m, Pcg& rng, Sample& s) {
  if (wo.z <= 0.f) return false;
  const float first = rng.next_f32();
  const float second = rng.next_f32();
  const F3 h = vndf_sample(wo, m.ax, m.ay, second, first);
  if (h.z <= 0.f) return false;
  s.wi = reflect(-wo, h);
  if (s.wi.z <= 0.f) return false;
  s.pdf = ggxr_pdf(wo, s.wi, m);
  if (s.pdf < 1e-8f) return false;
  s.f = ggxr_f(wo, s.wi, m);
  s.delta = false;
  return true;
}

// GGX transmission, Walter 2007 (bsdf.h:364-460)
inline F3 ggxt_f(F3 wo, F3 wi, const Mat& m) {
  if (wo.z == 0.f || wi.z == 0.f) return F3::splat(0.f);
  if (wo.z * wi.z > 0.f) return F3::splat(0.f);
  const float eta = (wo.z > 0.f) ? (1.f / m.ior) : m.ior;
  F3 h = normalize(wo + eta * wi);
  if (h.z < 0.f) h = -h;
  const float oh = dot(wo, h), ih = dot(wi, h);
  if (oh * ih >= 0.f) return F3::splat(0.f);
  const float D = ggx_D(h, m.ax, m.ay);
  const float G = ggx_G2(wo, wi, m.ax, m.ay);
  const float F = fr_dielectric(oh, m.ior);
  const float co = std::fabs(wo.z), ci = std::fabs(wi.z);
  if (co < 1e-8f || ci < 1e-8f) return F3::splat(0.f);
  const float den = oh + eta * ih;
  const float den2 = den * den;
  if (std::fabs(den) < 1e-8f) return F3::splat(0.f);
  const float fac = std::fabs(oh * ih) * (eta * eta) / (co * ci * den2);
  return m.base * ((1.f - F) * D * G * fac);
}
inline float ggxt_pdf(F3 wo, F3 wi, const Mat& m) {
  if (wo.z == 0.f || wi.z == 0.f) return 0.f;
  if (wo.z * wi.z > 0.f) return 0.f;
  const float eta = (wo.z > 0.f) ? (1.f / m.ior) : m.ior;
  F3 h = normalize(wo + eta * wi);
  if (h.z < 0.f) h = -h;
  const float oh = dot(wo, h), ih = dot(wi, h);
  if (oh * ih >= 0.f) return 0.f;
  const float ph = vndf_pdf(wo, h, m.ax, m.ay);
  const float den = oh + eta * ih;
  if (std::fabs(den) < 1e-8f) return 0.f;
  const float jac = std::fabs((eta * eta * ih) / (den * den));
  return ph * jac;
}
inline bool ggxt_sample(F3 wo, const Mat& m, Pcg& rng, Sample& s) {
  if (std::fabs(wo.z) < 1e-8f) return false;
  const float first = rng.next_f32();
  const float second = rng.next_f32();
  F3 h = vndf_sample(wo, m.ax, m.ay, second, first);
  if (h.z < 0.f) h = -h;
  const float eta = (wo.z > 0.f) ? (1.f / m.ior) : m.ior;
  const float oh = dot(wo, h);
  const float s2o = stdmax(0.f, 1.f - oh * oh);
  const float s2i = eta * eta * s2o;
  if (s2i >= 1.f) return false;
  const float ci = std::sqrt(stdmax(0.f, 1.f - s2i));
  const float sg = (oh > 0.f) ? 1.f : -1.f;
  s.wi = normalize((-eta * wo) + (eta * oh - sg * ci) * h);
  if (wo.z * s.wi.z > 0.f) return false;
  s.pdf = ggxt_pdf(wo, s.wi, m);
  if (s.pdf < 1e-8f) return false;
  s.f = ggxt_f(wo, s.wi, m);
  s.delta = false;
  return true;
}

// perfect mirror (bsdf.h:468-507)
inline bool mirror_sample(F3 wo, const Mat& m, Sample& s) {
  if (std::fabs(wo.z) < 1e-8f) return false;
  const F3 n(0.f, 0.f, (wo.z > 0.f) ? 1.f : -1.f);
  s.wi = reflect(-wo, n);
  const F3 F = m.conductor ? fr_schlick(std::fabs(wo.z), m.f0)
                           : F3::splat(fr_dielectric(std::fabs(wo.z), m.ior));
  const float ci = std::fabs(s.wi.z);
  if (ci < 1e-8f) return false;
  s.f = F / ci;
  s.pdf = 1.f;
  s.delta = true;
  return true;
}

// perfect refraction (bsdf.h:515-548)
inline bool glass_sample(F3 wo, const Mat& m, Sample& s) {
  if (std::fabs(wo.z) < 1e-8f) return false;
  const float eta = (wo.z > 0.f) ? (1.f / m.ior) : m.ior;
  const F3 n(0.f, 0.f, (wo.z > 0.f) ? 1.f : -1.f);
  if (!refract_dir(-wo, n, eta, s.wi)) return false;
  s.wi = normalize(s.wi);
  const float F = fr_dielectric(std::fabs(wo.z), m.ior);
  const float ci = std::fabs(s.wi.z);
  if (ci < 1e-8f) return false;
  s.f = m.base * ((1.f - F) * eta * eta / ci);
  s.pdf = 1.f;
  s.delta = true;
  return true;
}

// diffuse-like subsurface lobe (bsdf.h:557-590)
inline F3 sss_f(F3 wi, const Mat& m) {
  if (wi.z <= 0.f) return F3::splat(0.f);
  return lerp(m.base, m.ss_color, m.subsurface) * (1.f - m.metallic) * kInvPi;
}
inline bool sss_sample(F3 wo, const Mat& m, Pcg& rng, Sample& s) {
  const float first = rng.next_f32();
  const float second = rng.next_f32();
  s.wi = cos_hemi(second, first);
  if (wo.z < 0.f) s.wi.z = -s.wi.z;
  s.pdf = std::fabs(s.wi.z) * kInvPi;
  if (s.pdf < 1e-8f) return false;
  s.f = lerp(m.base, m.ss_color, m.subsurface) * (1.f - m.metallic) * kInvPi;
  s.delta = false;
  return true;
}

// ---- NEE evaluation (bsdf.h:601-657) ----------------------------------------
inline F3 nee_f(F3 wo, F3 wi, const Mat& m) {
  if (m.delta) return F3::splat(0.f);
  if (m.conductor) return ggxr_f(wo, wi, m);
  if (m.transmissive) return (wi.z > 0.f) ? ggxr_f(wo, wi, m) : ggxt_f(wo, wi, m);
  if (m.sss) return sss_f(wi, m);
  const float ws = m.f0.x;
  const float wd = 1.f - ws;
  if (wi.z > 0.f) return lambert_f(wi, m) * wd + ggxr_f(wo, wi, m) * ws;
  return F3::splat(0.f);
}
inline float nee_pdf(F3 wo, F3 wi, const Mat& m) {
  if (m.delta) return 0.f;
  if (m.conductor) return ggxr_pdf(wo, wi, m);
  if (m.transmissive) {
    const float F = fr_dielectric(std::fabs(wo.z), m.ior);
    if (wi.z > 0.f) return F * ggxr_pdf(wo, wi, m);
    return (1.f - F) * ggxt_pdf(wo, wi, m);
  }
  if (m.sss) return lambert_pdf(wi);
  const float ws = m.f0.x;
  const float wd = 1.f - ws;
  if (wi.z > 0.f) return lambert_pdf(wi) * wd + ggxr_pdf(wo, wi, m) * ws;
  return 0.f;
}

// ---- lobe selection + sampling (bsdf.h:666-708) ------------------------------
inline bool sample_bsdf(F3 wo, const Mat& m, Pcg& rng, Sample& s) {
  if (m.delta) {
    if (m.transmissive) {
      const float F = fr_dielectric(wo.z, m.ior);
      if (rng.next_f32() < F) return mirror_sample(wo, m, s);
      return glass_sample(wo, m, s);
    }
    return mirror_sample(wo, m, s);
  }
  if (m.conductor) return ggxr_sample(wo, m, rng, s);
  if (m.sss) return sss_sample(wo, m, rng, s);
  if (m.transmissive) {
    const float F = fr_dielectric(std::fabs(wo.z), m.ior);
    if (rng.next_f32() < F) {
      if (!ggxr_sample(wo, m, rng, s)) return false;
      s.pdf *= F;
      return true;
    }
    if (!ggxt_sample(wo, m, rng, s)) return false;
    s.pdf *= (1.f - F);
    return true;
  }
  const float ws = stdclamp(m.f0.x, 0.f, 1.f);
  const float wd = 1.f - ws;
  if (rng.next_f32() < ws) {
    if (!ggxr_sample(wo, m, rng, s)) return false;
    s.pdf *= ws;
    return true;
  }
  if (!lambert_sample(wo, m, rng, s)) return false;
  s.pdf *= wd;
  return true;
}

}  // namespace orc
