// UsdPreviewSurface's closure on the device: a Fresnel-attenuated Lambert base
// and a GGX lobe, plus a share of rough or smooth dielectric interface.
// Behaviour follows backend/vulkan/shaders/bsdf/preview.slang (functions cited
// by line); directions are in the shading frame with z along the shading
// normal on wo's side.
#pragma once

#include "device_common.hpp"

namespace lotus {

constexpr float kMinAlpha = 1.0e-3f;  // preview.slang:17

struct Closure {
  V3 base_color;
  float alpha;
  float metallic;
  V3 dielectric_f0;
  float dielectric_f90;
  V3 f0;
  float f90;
  float transmission;
  V3 transmission_color;
  float eta;
  bool thin_walled;
  bool smooth;
};

// preview.slang:60-89.
__device__ __forceinline__ Closure make_preview_closure(V3 base_color, float roughness, float metallic,
                                                        float ior, bool specular_workflow, V3 specular,
                                                        float transmission, V3 transmission_color,
                                                        bool thin_walled, bool entering) {
  Closure b;
  b.base_color = base_color;
  b.alpha = maxf(roughness * roughness, kMinAlpha);
  b.metallic = specular_workflow ? 0.0f : metallic;
  b.transmission = (1.0f - b.metallic) * transmission;
  b.transmission_color = transmission_color;
  b.eta = thin_walled || entering ? ior : 1.0f / ior;
  b.thin_walled = thin_walled;
  b.smooth = roughness * roughness < kMinAlpha || ior == 1.0f;
  if (b.transmission > 0.0f)
    b.metallic = b.transmission < 1.0f ? minf(b.metallic / (1.0f - b.transmission), 1.0f) : 0.0f;
  const float ratio = 1.0f - 2.0f / (ior + 1.0f);
  const float r2 = ratio * ratio;
  b.dielectric_f0 = specular_workflow ? specular : V3{r2, r2, r2};
  b.dielectric_f90 = specular_workflow || ior > 1.0f ? 1.0f : 0.0f;
  b.f0 = base_color * b.metallic + b.dielectric_f0 * (1.0f - b.metallic);
  b.f90 = b.metallic + (1.0f - b.metallic) * b.dielectric_f90;
  return b;
}

__device__ __forceinline__ float pow5(float x) {
  const float x2 = x * x;
  return (x2 * x2) * x;
}
// preview.slang:92-95.
__device__ __forceinline__ V3 fresnel_schlick(V3 f0, float f90, float cos_theta) {
  const float w = pow5(saturate(1.0f - cos_theta));
  const float g = f90 * w;
  return f0 * (1.0f - w) + V3{g, g, g};
}
// preview.slang:100-104.
__device__ __forceinline__ V3 dielectric_transmission(const Closure& b, float cos_theta) {
  const V3 fr = fresnel_schlick(b.dielectric_f0, b.dielectric_f90, cos_theta);
  return V3{1.0f, 1.0f, 1.0f} + fr * -1.0f;
}
// preview.slang:106-112.
__device__ __forceinline__ float specular_probability(const Closure& b, V3 wo) {
  if (b.metallic == 1.0f) return 1.0f;
  if (b.dielectric_f90 == 0.0f) return b.metallic;
  const float d = minf(maxf(max3(fresnel_schlick(b.dielectric_f0, b.dielectric_f90, wo.z)), 0.05f), 0.95f);
  return b.metallic + (1.0f - b.metallic) * d;
}
// preview.slang:118-127.
__device__ __forceinline__ float fresnel_dielectric(float cos_theta, float eta) {
  cos_theta = saturate(cos_theta);
  const float sin2_t = (1.0f - cos_theta * cos_theta) / (eta * eta);
  if (sin2_t >= 1.0f) return 1.0f;
  const float cos_t = sqrtf(1.0f - sin2_t);
  const float par = (eta * cos_theta - cos_t) / (eta * cos_theta + cos_t);
  const float per = (cos_theta - eta * cos_t) / (cos_theta + eta * cos_t);
  return 0.5f * (par * par + per * per);
}
// preview.slang:132-135.
__device__ __forceinline__ float interface_reflectance(const Closure& b, float cos_theta) {
  const float r = fresnel_dielectric(cos_theta, b.eta);
  return b.thin_walled ? 2.0f * r / (1.0f + r) : r;
}
// preview.slang:138-148.
__device__ __forceinline__ float ggx_d(V3 h, float alpha) {
  const float a2 = alpha * alpha;
  const float d = h.z * h.z * (a2 - 1.0f) + 1.0f;
  return a2 / (kPi * d * d);
}
__device__ __forceinline__ float ggx_lambda(V3 w, float alpha) {
  const float tan2 = maxf(0.0f, 1.0f - w.z * w.z) / (w.z * w.z);
  return 0.5f * (sqrtf(1.0f + alpha * alpha * tan2) - 1.0f);
}
// preview.slang:152-160 (Dupuy and Benyoub 2023, spherical caps).
__device__ __forceinline__ V3 sample_ggx_visible_normal(V3 wo, float alpha, float ux, float uy) {
  const V3 wo_std = normalize3(V3{wo.x * alpha, wo.y * alpha, wo.z});
  float s, c;
  sincos_turns(ux, &s, &c);
  const float z = (1.0f - uy) * (1.0f + wo_std.z) - wo_std.z;
  const float sin_theta = sqrtf(saturate(1.0f - z * z));
  const V3 h_std = V3{sin_theta * c, sin_theta * s, z} + wo_std;
  return normalize3(V3{h_std.x * alpha, h_std.y * alpha, h_std.z});
}
// preview.slang:164-183.
__device__ __forceinline__ V3 ggx_eval_cos(const Closure& b, V3 wo, V3 wi) {
  const V3 h = normalize3(wo + wi);
  const float g2 = 1.0f / (1.0f + ggx_lambda(wo, b.alpha) + ggx_lambda(wi, b.alpha));
  return fresnel_schlick(b.f0, b.f90, dot3(wo, h)) * (ggx_d(h, b.alpha) * g2 / (4.0f * wo.z));
}
__device__ __forceinline__ float ggx_pdf(const Closure& b, V3 wo, V3 wi) {
  const V3 h = normalize3(wo + wi);
  const float g1 = 1.0f / (1.0f + ggx_lambda(wo, b.alpha));
  return ggx_d(h, b.alpha) * g1 / (4.0f * wo.z);
}
__device__ __forceinline__ float smith_ratio(const Closure& b, V3 wo, V3 wi) {
  const float lo = ggx_lambda(wo, b.alpha);
  return (1.0f + lo) / (1.0f + lo + ggx_lambda(wi, b.alpha));
}
// preview.slang:187-212.
__device__ __forceinline__ V3 layered_eval_cos(const Closure& b, V3 wo, V3 wi) {
  if (wo.z <= 0.0f || wi.z <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
  V3 v = b.base_color * ((1.0f - b.metallic) * wi.z / kPi);
  v = v * dielectric_transmission(b, wo.z) * dielectric_transmission(b, wi.z);
  if (b.f90 > 0.0f) v = v + ggx_eval_cos(b, wo, wi);
  return v;
}
__device__ __forceinline__ float layered_pdf(const Closure& b, V3 wo, V3 wi) {
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  const float spec = specular_probability(b, wo);
  float pdf = (1.0f - spec) * wi.z / kPi;
  if (spec > 0.0f) pdf += spec * ggx_pdf(b, wo, wi);
  return pdf;
}
// preview.slang:216-252.
__device__ __forceinline__ float interface_reflection_eval_cos(const Closure& b, V3 wo, V3 wi) {
  const V3 h = normalize3(wo + wi);
  const float g2 = 1.0f / (1.0f + ggx_lambda(wo, b.alpha) + ggx_lambda(wi, b.alpha));
  return interface_reflectance(b, dot3(wo, h)) * ggx_d(h, b.alpha) * g2 / (4.0f * wo.z);
}
__device__ __forceinline__ V3 reflection_evaluate(const Closure& b, V3 wo, V3 wi) {
  const V3 layered = layered_eval_cos(b, wo, wi);
  if (b.transmission == 0.0f) return layered;
  V3 v = layered * (1.0f - b.transmission);
  if (!b.smooth && wo.z > 0.0f && wi.z > 0.0f) {
    const float r = b.transmission * interface_reflection_eval_cos(b, wo, wi);
    v = v + V3{r, r, r};
  }
  return v;
}
__device__ __forceinline__ float reflection_pdf(const Closure& b, V3 wo, V3 wi) {
  const float layered = layered_pdf(b, wo, wi);
  if (b.transmission == 0.0f) return layered;
  float pdf = (1.0f - b.transmission) * layered;
  if (!b.smooth && wo.z > 0.0f && wi.z > 0.0f) {
    const V3 h = normalize3(wo + wi);
    pdf += b.transmission * interface_reflectance(b, dot3(wo, h)) * ggx_pdf(b, wo, wi);
  }
  return pdf;
}
// preview.slang:262-310.
struct Transmit {
  float value, pdf;
};
__device__ __forceinline__ Transmit preview_transmission(const Closure& b, V3 wo, V3 wi) {
  Transmit r{0.0f, 0.0f};
  if (b.transmission == 0.0f || b.smooth || !(wo.z > 0.0f) || !(wi.z < 0.0f)) return r;
  if (b.thin_walled) {
    const V3 refl = V3{wi.x, wi.y, -wi.z};
    const V3 h = normalize3(wo + refl);
    const float share = b.transmission * (1.0f - interface_reflectance(b, dot3(wo, h)));
    const float g2 = 1.0f / (1.0f + ggx_lambda(wo, b.alpha) + ggx_lambda(refl, b.alpha));
    r.value = share * ggx_d(h, b.alpha) * g2 / (4.0f * wo.z);
    r.pdf = share * ggx_pdf(b, wo, refl);
    return r;
  }
  V3 h = normalize3(wo + b.eta * wi);
  if (h.z < 0.0f) h = -h;
  const float cos_oh = dot3(wo, h);
  const float cos_ih = dot3(wi, h);
  if (!(cos_oh > 0.0f) || !(cos_ih < 0.0f)) return r;
  const float den = cos_ih + cos_oh / b.eta;
  const float jac = -cos_ih / (den * den);
  const float share = b.transmission * (1.0f - interface_reflectance(b, cos_oh));
  const float d = ggx_d(h, b.alpha);
  const float lo = ggx_lambda(wo, b.alpha);
  const float g1 = 1.0f / (1.0f + lo);
  const float g2 = 1.0f / (1.0f + lo + ggx_lambda(wi, b.alpha));
  r.pdf = share * (g1 * cos_oh * d / wo.z) * jac;
  r.value = share * (g2 * cos_oh * d / wo.z) * jac / (b.eta * b.eta);
  return r;
}
// preview.slang:314-327 (BsdfEvaluate / BsdfPdf, bsdf.slang:54-76).
__device__ __forceinline__ V3 bsdf_evaluate(const Closure& b, V3 wo, V3 wi) {
  if (wi.z < 0.0f) return b.transmission_color * preview_transmission(b, wo, wi).value;
  return reflection_evaluate(b, wo, wi);
}
__device__ __forceinline__ float bsdf_pdf(const Closure& b, V3 wo, V3 wi) {
  if (wi.z < 0.0f) return preview_transmission(b, wo, wi).pdf;
  return reflection_pdf(b, wo, wi);
}

struct BsdfSample {
  V3 wi;
  V3 weight;
  float pdf;
  bool delta;
  bool transmitted;
};
// preview.slang:348-405.
__device__ __forceinline__ bool sample_interface(const Closure& b, V3 wo, float ux, float uy, float uz,
                                                 BsdfSample& s) {
  const V3 h = b.smooth ? V3{0.0f, 0.0f, 1.0f} : sample_ggx_visible_normal(wo, b.alpha, ux, uy);
  const float cos_oh = dot3(wo, h);
  if (!(cos_oh > 0.0f)) return false;
  const V3 reflected = (2.0f * cos_oh) * h - wo;
  s.delta = b.smooth;
  if (uz < interface_reflectance(b, cos_oh)) {
    s.wi = reflected;
    if (b.smooth) {
      s.weight = V3{1.0f, 1.0f, 1.0f};
      return true;
    }
    if (reflected.z <= 0.0f) return false;
    const float r = smith_ratio(b, wo, reflected);
    s.weight = V3{r, r, r};
    s.pdf = reflection_pdf(b, wo, reflected);
    return s.pdf > 0.0f && finite3(s.weight);
  }
  s.transmitted = true;
  if (b.thin_walled) {
    if (!b.smooth && reflected.z <= 0.0f) return false;
    s.wi = V3{reflected.x, reflected.y, -reflected.z};
    s.weight = b.transmission_color * (b.smooth ? 1.0f : smith_ratio(b, wo, reflected));
    if (!b.smooth) s.pdf = preview_transmission(b, wo, s.wi).pdf;
    return finite3(s.weight);
  }
  const float sin2_t = (1.0f - cos_oh * cos_oh) / (b.eta * b.eta);
  if (sin2_t >= 1.0f) return false;
  const float cos_t = sqrtf(1.0f - sin2_t);
  s.wi = -wo / b.eta + (cos_oh / b.eta - cos_t) * h;
  if (s.wi.z >= 0.0f) return false;
  s.weight = b.transmission_color * ((b.smooth ? 1.0f : smith_ratio(b, wo, s.wi)) / (b.eta * b.eta));
  if (!b.smooth) s.pdf = preview_transmission(b, wo, s.wi).pdf;
  return finite3(s.weight);
}
// preview.slang:411-454 (BsdfSample, bsdf.slang:81-90).
__device__ __forceinline__ bool bsdf_sample(const Closure& b, V3 wo, float ux, float uy, float uz,
                                            BsdfSample& s) {
  s.wi = V3{0.0f, 0.0f, 0.0f};
  s.weight = V3{0.0f, 0.0f, 0.0f};
  s.pdf = 0.0f;
  s.delta = false;
  s.transmitted = false;
  if (wo.z <= 0.0f) return false;
  if (b.transmission > 0.0f) {
    if (uz < b.transmission) return sample_interface(b, wo, ux, uy, uz / b.transmission, s);
    uz = (uz - b.transmission) / (1.0f - b.transmission);
  }
  if (uz < specular_probability(b, wo)) {
    const V3 h = sample_ggx_visible_normal(wo, b.alpha, ux, uy);
    s.wi = (2.0f * dot3(wo, h)) * h - wo;
  } else {
    // sampling.slang:79-84
    const float radius = sqrtf(ux);
    float sn, cs;
    sincos_turns(uy, &sn, &cs);
    s.wi = V3{radius * cs, radius * sn, sqrtf(maxf(0.0f, 1.0f - ux))};
  }
  const float lpdf = layered_pdf(b, wo, s.wi);
  if (!(lpdf > 0.0f)) return false;
  s.pdf = b.transmission > 0.0f ? reflection_pdf(b, wo, s.wi) : lpdf;
  s.weight = layered_eval_cos(b, wo, s.wi) / lpdf;
  if (b.metallic == 1.0f) {
    const V3 h = normalize3(wo + s.wi);
    const float lo = ggx_lambda(wo, b.alpha);
    s.weight = fresnel_schlick(b.f0, b.f90, dot3(wo, h)) *
               ((1.0f + lo) / (1.0f + lo + ggx_lambda(s.wi, b.alpha)));
  }
  return finite3(s.weight);
}

}  // namespace lotus
