// glTF PBR's closure on the device (MaterialX's gltf_pbr surface):
// directional-albedo table lookups (backend/vulkan/shaders/bsdf/albedo.slang),
// the primitive lobes it stacks (lobes.slang), the model's terms and its
// composed closure (gltf_pbr.slang, composed.slang).
//
// The reference's composed closure is a list of up to six lobes that the
// model's builder appends and then mixes and layers. glTF PBR always builds
// the same list in the same order, so here each possible lobe has a fixed
// slot (coat, sheen, dielectric reflection, diffuse, thin transmission,
// metal) and an absent lobe is skipped: the operations touch the same lobes
// in the same order with the same float operations, and sums over the slots
// add only the present lobes, as the list's sums do. Directions are in the
// shading frame, z along the shading normal on wo's side.
#pragma once

#include "bsdf.hpp"

namespace lotus {

// --- portable transcendental functions ---------------------------------------
// Shared by the kernels and the oracle, so both round alike: the natural
// logarithm by the atanh series of the mantissa reduced to [sqrt(1/2),
// sqrt(2)) (|s| <= 0.172, terms to s^9), and exp by 2^k times a degree-7
// Taylor polynomial of the remainder, ln 2 split in two.
__device__ __forceinline__ float log_positive(float x) {  // x > 0, finite
  int e;
  float m = frexpf(x, &e);
  if (m < 0.70710678f) {
    m = m * 2.0f;
    e -= 1;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  const float p = s2 * (0.333333343f + s2 * (0.2f + s2 * (0.142857149f + s2 * 0.111111112f)));
  return static_cast<float>(e) * 0.693147182f + (2.0f * s + (2.0f * s) * p);
}
// e^t for t <= 0; 0 below e^-87, under the smallest normal float.
__device__ __forceinline__ float exp_nonpositive(float t) {
  if (!(t > -87.0f)) return 0.0f;
  const float k = floorf(t * 1.44269504f + 0.5f);
  const float r = (t - k * 0.693145752f) - k * 1.42860677e-06f;
  const float p =
      1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (0.0416666679f + r * (0.00833333377f +
                                                                                   r * (0.00138888892f + r * 0.000198412701f))))));
  return ldexpf(p, static_cast<int>(k));
}
// x^y for x in [0, 1] and y > 0.
__device__ __forceinline__ float pow_unit(float x, float y) {
  if (!(x > 0.0f)) return 0.0f;
  if (x >= 1.0f) return 1.0f;
  return exp_nonpositive(y * log_positive(x));
}

// --- albedo tables (albedo.slang:22-373) ------------------------------------
constexpr uint32_t kGgxCosines = 32, kDielectricCosines = 48, kAlbedoAlphas = 32, kAlbedoIndices = 16;
constexpr uint32_t kTableGgx = 0, kTableSchlick = 1024, kTableDielectric = 2048, kTableSheen = 75776;
constexpr uint32_t kSheetLeaving = 1, kSheetThin = 2;
constexpr float kMinCosine = 2.0e-4f, kCosineSlope = 4.0f, kAlphaSlope = 2.0f, kIndexSlope = 0.5f;
constexpr float kMinSheenAxis = 0.005f, kAlbedoMinIndex = 1.01f, kAlbedoMaxIndex = 3.0f, kMinEvanescent = 2.0e-3f;
constexpr uint32_t kCriticalNode = kDielectricCosines / 2;
constexpr float kCriticalCoordinate = static_cast<float>(kCriticalNode) / static_cast<float>(kDielectricCosines - 1);

__device__ __forceinline__ float log_linear_scale(float v, float slope) { return log_positive(v) + slope * v; }
// albedo.slang:99-104.
__device__ __forceinline__ float log_linear_coordinate(float v, float low, float high, float slope) {
  const float c = minf(maxf(v, low), high);
  const float l = log_linear_scale(low, slope);
  return (log_linear_scale(c, slope) - l) / (log_linear_scale(high, slope) - l);
}
__device__ __forceinline__ float cosine_coordinate(float c) { return log_linear_coordinate(c, kMinCosine, 1.0f, kCosineSlope); }
__device__ __forceinline__ float alpha_coordinate(float a) { return log_linear_coordinate(a, kMinAlpha, 1.0f, kAlphaSlope); }
__device__ __forceinline__ float sheen_coordinate(float r) { return log_linear_coordinate(r, kMinSheenAxis, 1.0f, kAlphaSlope); }
__device__ __forceinline__ float index_coordinate(float i) {
  return log_linear_coordinate(i - 1.0f, kAlbedoMinIndex - 1.0f, kAlbedoMaxIndex - 1.0f, kIndexSlope);
}
// albedo.slang:162-198 (CriticalCosine, EvanescentScale, LeavingCoordinate).
__device__ __forceinline__ float critical_cosine(float index) { return sqrtf(maxf(1.0f - 1.0f / (index * index), 0.0f)); }
__device__ __forceinline__ float evanescent_scale(float mu, float index, float critical) {
  const float tau = maxf(index * sqrtf(maxf(critical * critical - mu * mu, 0.0f)), kMinEvanescent);
  return (log_positive(mu) - log_positive(tau)) + kCosineSlope * mu;
}
__device__ __forceinline__ float leaving_coordinate(float cos_theta, float index) {
  const float critical = critical_cosine(index);
  if (cos_theta < critical) {
    const float low = evanescent_scale(kMinCosine, index, critical);
    const float high = evanescent_scale(critical, index, critical);
    return kCriticalCoordinate * saturate((evanescent_scale(maxf(cos_theta, kMinCosine), index, critical) - low) /
                                          (high - low));
  }
  const float refracted = sqrtf(maxf(1.0f - index * index * (1.0f - cos_theta * cos_theta), 0.0f));
  return kCriticalCoordinate + cosine_coordinate(refracted) * (1.0f - kCriticalCoordinate);
}

// Catmull-Rom weights (albedo.slang:283-289).
__device__ __forceinline__ void cubic_weights(float t, float w[4]) {
  const float t2 = t * t, t3 = t2 * t;
  w[0] = ((-0.5f * t3) + t2) - 0.5f * t;
  w[1] = ((1.5f * t3) - 2.5f * t2) + 1.0f;
  w[2] = ((-1.5f * t3) + 2.0f * t2) + 0.5f * t;
  w[3] = (0.5f * t3) - 0.5f * t2;
}
// A node of a slice, extrapolated linearly past [first, last] and the alpha
// axis (albedo.slang:294-309).
__device__ __forceinline__ float fetch_node(const float* T, uint32_t base, int cosines, int c, int a, int first,
                                            int last) {
  const int last_alpha = static_cast<int>(kAlbedoAlphas) - 1;
  const int cc = c < first ? first : (c > last ? last : c);
  const int ac = a < 0 ? 0 : (a > last_alpha ? last_alpha : a);
  const int dc = c < first ? 1 : (c > last ? -1 : 0);
  const int da = a < 0 ? 1 : (a > last_alpha ? -1 : 0);
  const float at = T[base + static_cast<uint32_t>(ac * cosines + cc)];
  if (dc == 0 && da == 0) return at;
  const float inner = T[base + static_cast<uint32_t>((ac + da) * cosines + cc + dc)];
  return 2.0f * at - inner;
}
// albedo.slang:311-337.
__device__ __forceinline__ float fetch_slice(const float* T, uint32_t base, int cosines, float x_cosine, float x_alpha,
                                             bool split) {
  const float fc = saturate(x_cosine) * static_cast<float>(cosines - 1);
  const float fa = saturate(x_alpha) * static_cast<float>(kAlbedoAlphas - 1);
  const uint32_t uc = static_cast<uint32_t>(fc), ua = static_cast<uint32_t>(fa);
  const int c0 = static_cast<int>(uc < static_cast<uint32_t>(cosines - 2) ? uc : static_cast<uint32_t>(cosines - 2));
  const int a0 = static_cast<int>(ua < kAlbedoAlphas - 2 ? ua : kAlbedoAlphas - 2);
  int first = 0, last = cosines - 1;
  if (split) {
    if (c0 < static_cast<int>(kCriticalNode)) last = static_cast<int>(kCriticalNode);
    else first = static_cast<int>(kCriticalNode);
  }
  float wc[4], wa[4];
  cubic_weights(fc - static_cast<float>(c0), wc);
  cubic_weights(fa - static_cast<float>(a0), wa);
  float value = 0.0f;
  for (int j = 0; j < 4; ++j) {
    float row = 0.0f;
    for (int i = 0; i < 4; ++i) row = row + wc[i] * fetch_node(T, base, cosines, c0 + i - 1, a0 + j - 1, first, last);
    value = value + wa[j] * row;
  }
  return value;
}
// albedo.slang:339-371.
__device__ __forceinline__ float albedo_ggx(const float* T, float c, float alpha) {
  return fetch_slice(T, kTableGgx, kGgxCosines, cosine_coordinate(c), alpha_coordinate(alpha), false);
}
__device__ __forceinline__ float albedo_schlick_share(const float* T, float c, float alpha) {
  return fetch_slice(T, kTableSchlick, kGgxCosines, cosine_coordinate(c), alpha_coordinate(alpha), false);
}
__device__ __forceinline__ float albedo_sheen(const float* T, float c, float roughness) {
  return maxf(fetch_slice(T, kTableSheen, kGgxCosines, cosine_coordinate(c), sheen_coordinate(roughness), false), 0.0f);
}
// A rough dielectric's reflected albedo (albedo.slang:373-410); a thin wall's
// index below 1 is looked up as its reciprocal.
__device__ __forceinline__ float albedo_dielectric(const float* T, float c, float alpha, float eta, bool thin) {
  uint32_t sheet = thin ? kSheetThin : 0u;
  float index = eta;
  if (eta < 1.0f) {
    index = 1.0f / eta;
    sheet = thin ? kSheetThin : kSheetLeaving;
  }
  index = minf(maxf(index, kAlbedoMinIndex), kAlbedoMaxIndex);
  const float fi = index_coordinate(index) * static_cast<float>(kAlbedoIndices - 1);
  const uint32_t ui = static_cast<uint32_t>(fi);
  const int i0 = static_cast<int>(ui < kAlbedoIndices - 2 ? ui : kAlbedoIndices - 2);
  const bool split = sheet == kSheetLeaving;
  const float x_cosine = split ? leaving_coordinate(c, index) : cosine_coordinate(c);
  const float x_alpha = alpha_coordinate(alpha);
  const uint32_t slice_size = kDielectricCosines * kAlbedoAlphas;
  const uint32_t base = kTableDielectric + sheet * kAlbedoIndices * slice_size;
  float slices[4];
  for (int j = 0; j < 4; ++j) {
    int slice = i0 + j - 1;
    slice = slice < 0 ? 0 : (slice > static_cast<int>(kAlbedoIndices) - 1 ? static_cast<int>(kAlbedoIndices) - 1 : slice);
    slices[j] = fetch_slice(T, base + static_cast<uint32_t>(slice) * slice_size, kDielectricCosines, x_cosine, x_alpha,
                            split);
  }
  if (i0 == 0) slices[0] = 2.0f * slices[1] - slices[2];
  if (i0 == static_cast<int>(kAlbedoIndices) - 2) slices[3] = 2.0f * slices[2] - slices[1];
  float w[4];
  cubic_weights(fi - static_cast<float>(i0), w);
  return ((w[0] * slices[0] + w[1] * slices[1]) + w[2] * slices[2]) + w[3] * slices[3];
}

// --- lobes (lobes.slang) -------------------------------------------------------
constexpr float kGltfClearcoatIndex = 1.5f;  // gltf_pbr.slang:29
constexpr float kMinSheenRoughness = 0.005f;  // lobes.slang:37
constexpr float kSelectionFloor = 0.05f;      // composed.slang:41

// An orthonormal basis about a unit normal (Duff et al. 2017), lobes.slang:141-152.
struct LobeFrame {
  V3 t, b, n;
};
__device__ __forceinline__ LobeFrame lobe_frame(V3 n) {
  const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sign + n.z);
  const float b = (n.x * n.y) * a;
  return LobeFrame{V3{1.0f + ((sign * n.x) * n.x) * a, sign * b, -sign * n.x}, V3{b, sign + (n.y * n.y) * a, -n.y}, n};
}
__device__ __forceinline__ V3 to_lobe(const LobeFrame& f, V3 w) { return V3{dot3(w, f.t), dot3(w, f.b), dot3(w, f.n)}; }
__device__ __forceinline__ V3 from_lobe(const LobeFrame& f, V3 w) { return ((w.x * f.t) + (w.y * f.b)) + (w.z * f.n); }

// GGX reflection's D G2 / (4 cos_o) and its visible-normal density (lobes.slang:185-196).
__device__ __forceinline__ float ggx_reflection(V3 wo, V3 wi, V3 h, float alpha) {
  const float g2 = 1.0f / ((1.0f + ggx_lambda(wo, alpha)) + ggx_lambda(wi, alpha));
  return ggx_d(h, alpha) * g2 / (4.0f * wo.z);
}
__device__ __forceinline__ float ggx_reflection_pdf(V3 wo, V3 h, float alpha) {
  const float g1 = 1.0f / (1.0f + ggx_lambda(wo, alpha));
  return ggx_d(h, alpha) * g1 / (4.0f * wo.z);
}
// Schlick with colour f90 (lobes.slang:168-171).
__device__ __forceinline__ V3 schlick_spectrum(V3 f0, V3 f90, float c) {
  const float w = pow5(saturate(1.0f - c));
  return f0 * (1.0f - w) + f90 * w;
}

// The glTF closure's lobes in their slots.
enum GltfSlot : int { kSlotCoat = 0, kSlotSheen = 1, kSlotTop = 2, kSlotDiffuse = 3, kSlotTransmit = 4, kSlotMetal = 5 };
constexpr int kGltfSlots = 6;

// What a glTF closure is built from (gltf_pbr.slang:57-97): the terms a
// stored reservoir surface keeps, 22 floats.
struct GltfTerms {
  V3 base_color;
  float alpha, metallic, e1, share;
  V3 f0;
  float f90, transmission, coat_scale, coat_dim, coat_alpha;
  V3 coat_normal;
  V3 sheen_color;
  float sheen_roughness, sheen_dim;
};

// The composed closure of glTF PBR for the outgoing direction it was built
// for: per slot its weight and selection probability (zero probability and
// weight for an absent lobe) and the compensations of the three GGX-Schlick
// lobes, with the terms their lobes read.
struct GltfClosure {
  GltfTerms t;
  uint32_t present;  // bit per slot
  V3 weight[kGltfSlots];
  float probability[kGltfSlots];
  V3 top_compensation, transmit_compensation, metal_compensation;
};

// The inputs at a hit (gltf_pbr.slang:31-56).
struct GltfInputs {
  V3 base_color;
  float roughness, metallic, dielectric_f0, specular;
  V3 specular_color;
  float transmission, clearcoat, clearcoat_roughness;
  V3 clearcoat_normal;
  V3 sheen_color;
  float sheen_roughness;
};

// MakeGltfPbrTerms (gltf_pbr.slang:99-155).
__device__ __forceinline__ GltfTerms make_gltf_terms(const float* T, const GltfInputs& in, V3 wo) {
  GltfTerms g;
  g.base_color = in.base_color;
  g.alpha = maxf(in.roughness * in.roughness, kMinAlpha);
  g.metallic = in.metallic;
  g.e1 = 0.0f;
  g.share = 0.0f;
  if (wo.z > 0.0f) {
    g.e1 = albedo_ggx(T, wo.z, g.alpha);
    g.share = albedo_schlick_share(T, wo.z, g.alpha);
  }
  const V3 f0 = in.specular_color * in.dielectric_f0;
  g.f0 = V3{saturate(f0.x), saturate(f0.y), saturate(f0.z)} * in.specular;
  g.f90 = in.specular;
  g.transmission = in.transmission;
  g.coat_scale = 0.0f;
  g.coat_dim = 0.0f;
  g.coat_alpha = maxf(in.clearcoat_roughness * in.clearcoat_roughness, kMinAlpha);
  g.coat_normal = in.clearcoat_normal.z > 0.0f ? in.clearcoat_normal : V3{0.0f, 0.0f, 1.0f};
  const float coat_cos = dot3(wo, g.coat_normal);
  if (in.clearcoat > 0.0f && coat_cos > 0.0f) {
    const float reflected = albedo_dielectric(T, coat_cos, g.coat_alpha, kGltfClearcoatIndex, false);
    const float coat_e1 = albedo_ggx(T, coat_cos, g.coat_alpha);
    const float root = (1.0f - kGltfClearcoatIndex) / (1.0f + kGltfClearcoatIndex);
    float compensation = 1.0f;
    if (coat_e1 > 0.0f) compensation = compensation + ((root * root) * (1.0f - coat_e1)) / coat_e1;
    g.coat_scale = in.clearcoat * compensation;
    g.coat_dim = saturate(g.coat_scale * reflected);
  }
  g.sheen_color = in.sheen_color;
  g.sheen_roughness = maxf(in.sheen_roughness * in.sheen_roughness, kMinSheenRoughness);
  g.sheen_dim = 0.0f;
  const float sheen = max3(in.sheen_color);
  if (sheen > 0.0f && wo.z > 0.0f) {
    const float e = albedo_sheen(T, wo.z, g.sheen_roughness);
    if (e > 1.0f) g.sheen_color = in.sheen_color * (1.0f / e);
    g.sheen_dim = saturate(sheen * minf(e, 1.0f));
  }
  return g;
}

// The lobes' values, densities and samples in their own frames
// (lobes.slang:271-508), for the slots' fixed lobe kinds.
__device__ __forceinline__ V3 gltf_lobe_value(const GltfClosure& c, int k, V3 wo, V3 wi) {
  const V3 zero{0.0f, 0.0f, 0.0f};
  if (!(wo.z > 0.0f)) return zero;
  const GltfTerms& t = c.t;
  if (k == kSlotDiffuse) return wi.z > 0.0f ? t.base_color * (wi.z / kPi) : zero;
  if (k == kSlotSheen) {
    if (!(wi.z > 0.0f)) return zero;
    const V3 h = normalize3(wo + wi);
    const float inverse = 1.0f / t.sheen_roughness;
    const float sin2 = maxf(1.0f - h.z * h.z, 0.0f);
    const float d = (2.0f + inverse) * pow_unit(sin2, 0.5f * inverse) / (2.0f * kPi);
    const float v = d * wi.z / (4.0f * ((wi.z + wo.z) - wi.z * wo.z));
    return V3{v, v, v};  // white colour, compensation 1
  }
  if (k == kSlotCoat) {
    if (!(wi.z > 0.0f)) return zero;  // the reflecting interface only
    const V3 h = normalize3(wo + wi);
    const float v = fresnel_dielectric(dot3(wo, h), kGltfClearcoatIndex) * ggx_reflection(wo, wi, h, t.coat_alpha);
    return V3{v, v, v};  // white tint, not compensated
  }
  // the GGX-Schlick lobes: the dielectric's reflection, the thin
  // transmission (the reflection mirrored through the tangent plane) and the
  // metal
  const V3 wr = k == kSlotTransmit ? V3{wi.x, wi.y, -wi.z} : wi;
  if (!(wr.z > 0.0f)) return zero;
  const V3 h = normalize3(wo + wr);
  const float g = ggx_reflection(wo, wr, h, t.alpha);
  const float ch = dot3(wo, h);
  if (k == kSlotTop) return (schlick_spectrum(t.f0, V3{t.f90, t.f90, t.f90}, ch) * c.top_compensation) * g;
  if (k == kSlotTransmit)
    return (schlick_spectrum(V3{1.0f, 1.0f, 1.0f}, V3{1.0f, 1.0f, 1.0f}, ch) * c.transmit_compensation) * g;
  return (schlick_spectrum(t.base_color, V3{1.0f, 1.0f, 1.0f}, ch) * c.metal_compensation) * g;
}

__device__ __forceinline__ float gltf_lobe_pdf(const GltfClosure& c, int k, V3 wo, V3 wi) {
  if (!(wo.z > 0.0f)) return 0.0f;
  if (k == kSlotDiffuse) return wi.z > 0.0f ? wi.z / kPi : 0.0f;
  if (k == kSlotSheen) return wi.z > 0.0f ? 1.0f / (2.0f * kPi) : 0.0f;
  if (k == kSlotCoat) return wi.z > 0.0f ? ggx_reflection_pdf(wo, normalize3(wo + wi), c.t.coat_alpha) : 0.0f;
  const V3 wr = k == kSlotTransmit ? V3{wi.x, wi.y, -wi.z} : wi;
  return wr.z > 0.0f ? ggx_reflection_pdf(wo, normalize3(wo + wr), c.t.alpha) : 0.0f;
}

__device__ __forceinline__ bool gltf_lobe_sample(const GltfClosure& c, int k, V3 wo, float ux, float uy, V3& wi) {
  if (!(wo.z > 0.0f)) return false;
  if (k == kSlotDiffuse) {
    // sampling.slang:79-84 (SampleCosineHemisphere)
    const float radius = sqrtf(ux);
    float sn, cs;
    sincos_turns(uy, &sn, &cs);
    wi = V3{radius * cs, radius * sn, sqrtf(maxf(0.0f, 1.0f - ux))};
    return wi.z > 0.0f;
  }
  if (k == kSlotSheen) {
    // uniform over the hemisphere
    const float sin_theta = sqrtf(maxf(1.0f - ux * ux, 0.0f));
    float sn, cs;
    sincos_turns(uy, &sn, &cs);
    wi = V3{sin_theta * cs, sin_theta * sn, ux};
    return wi.z > 0.0f;
  }
  const float alpha = k == kSlotCoat ? c.t.coat_alpha : c.t.alpha;
  const V3 h = sample_ggx_visible_normal(wo, alpha, ux, uy);
  if (k == kSlotCoat) {
    const float cos_oh = dot3(wo, h);
    if (!(cos_oh > 0.0f)) return false;
    wi = (2.0f * cos_oh) * h - wo;
    return wi.z > 0.0f;
  }
  wi = (2.0f * dot3(wo, h)) * h - wo;
  if (!(wi.z > 0.0f)) return false;
  if (k == kSlotTransmit) wi.z = -wi.z;
  return true;
}

// Whether slot k's lobe scatters into wi's hemisphere of the shading frame
// (lobes.slang:127-133): the thin transmission below, the others above.
__device__ __forceinline__ bool gltf_scatters_toward(int k, V3 wi) { return k == kSlotTransmit ? wi.z < 0.0f : wi.z > 0.0f; }

// The slot's lobe normal: the coat's own, the shading normal otherwise.
__device__ __forceinline__ V3 gltf_lobe_normal(const GltfClosure& c, int k) {
  return k == kSlotCoat ? c.t.coat_normal : V3{0.0f, 0.0f, 1.0f};
}

// MakeGltfPbrClosure (gltf_pbr.slang:218-276) through the composed closure's
// operations (composed.slang:113-357).
__device__ __forceinline__ GltfClosure make_gltf_closure(const GltfTerms& t, V3 wo) {
  GltfClosure c;
  c.t = t;
  c.present = 0u;
  V3 albedo[kGltfSlots], estimate[kGltfSlots];
  const V3 zero{0.0f, 0.0f, 0.0f}, one{1.0f, 1.0f, 1.0f};
  for (int k = 0; k < kGltfSlots; ++k) {
    c.weight[k] = zero;
    c.probability[k] = 0.0f;
    albedo[k] = zero;
    estimate[k] = zero;
  }
  c.top_compensation = one;
  c.transmit_compensation = one;
  c.metal_compensation = one;
  auto push = [&](int k, V3 weight, V3 a, V3 e) {
    c.present |= 1u << k;
    c.weight[k] = weight;
    albedo[k] = a;
    estimate[k] = e;
  };
  // ComposedAddSchlick: compensation and albedo about the z normal
  auto schlick = [&](V3 f0, V3 f90, bool transmit, V3& compensation, V3& a) {
    compensation = one;
    a = zero;
    if (wo.z > 0.0f) {  // dot(wo, (0, 0, 1))
      if (t.e1 > 0.0f) compensation = one + f0 * ((1.0f - t.e1) / t.e1);
      a = (f0 * (t.e1 - t.share) + f90 * t.share) * compensation;
    }
    (void)transmit;
  };
  auto scale = [&](int begin, int end, V3 s) {
    for (int k = begin; k < end; ++k)
      if (c.present & (1u << k)) c.weight[k] = c.weight[k] * s;
  };
  auto range_albedo = [&](int begin, int end) {
    V3 sum = zero;
    for (int k = begin; k < end; ++k)
      if (c.present & (1u << k)) sum = sum + c.weight[k] * albedo[k];
    return sum;
  };
  if (t.coat_scale > 0.0f) {
    const float a = t.coat_dim / t.coat_scale;
    push(kSlotCoat, V3{t.coat_scale, t.coat_scale, t.coat_scale}, V3{a, a, a}, V3{a, a, a});
  }
  if (t.sheen_dim > 0.0f) {
    const float a = t.sheen_dim / max3(t.sheen_color);
    push(kSlotSheen, t.sheen_color, V3{a, a, a}, V3{a, a, a});
  }
  const bool dielectric = t.metallic < 1.0f;
  if (dielectric && t.f90 > 0.0f) {
    V3 a;
    schlick(t.f0, V3{t.f90, t.f90, t.f90}, false, c.top_compensation, a);
    push(kSlotTop, one, a, a);
  }
  if (dielectric && t.transmission < 1.0f) {
    const V3 a = wo.z > 0.0f ? t.base_color : zero;
    push(kSlotDiffuse, one, a, a);
  }
  if (dielectric && t.transmission > 0.0f) {
    V3 a;
    schlick(one, one, true, c.transmit_compensation, a);
    push(kSlotTransmit, one, zero, a);  // a mirrored lobe reflects nothing a layer above loses
    scale(kSlotTransmit, kSlotTransmit + 1, t.base_color);
  }
  if (t.transmission > 0.0f) {
    scale(kSlotDiffuse, kSlotTransmit, V3{1.0f - t.transmission, 1.0f - t.transmission, 1.0f - t.transmission});
    scale(kSlotTransmit, kSlotMetal, V3{t.transmission, t.transmission, t.transmission});
  }
  {
    const float s = saturate(1.0f - max3(range_albedo(kSlotTop, kSlotDiffuse)));
    scale(kSlotDiffuse, kSlotMetal, V3{s, s, s});
  }
  if (t.metallic > 0.0f) {
    V3 a;
    schlick(t.base_color, one, false, c.metal_compensation, a);
    push(kSlotMetal, one, a, a);
  }
  scale(kSlotTop, kSlotMetal, V3{1.0f - t.metallic, 1.0f - t.metallic, 1.0f - t.metallic});
  scale(kSlotMetal, kGltfSlots, V3{t.metallic, t.metallic, t.metallic});
  {
    const float s = saturate(1.0f - max3(range_albedo(kSlotSheen, kSlotTop)));
    scale(kSlotTop, kGltfSlots, V3{s, s, s});
  }
  {
    const V3 top = range_albedo(kSlotCoat, kSlotSheen);
    scale(kSlotSheen, kGltfSlots, V3{saturate(1.0f - top.x), saturate(1.0f - top.y), saturate(1.0f - top.z)});
  }
  // ComposedFinalize (composed.slang:307-333)
  float total = 0.0f;
  for (int k = 0; k < kGltfSlots; ++k) {
    if (!(c.present & (1u << k))) continue;
    const float e = max3(c.weight[k] * estimate[k]);
    c.probability[k] = e > 0.0f ? e : 0.0f;
    total += c.probability[k];
  }
  if (!(total > 0.0f)) return c;
  float floored = 0.0f;
  for (int k = 0; k < kGltfSlots; ++k)
    if (c.probability[k] > 0.0f) {
      c.probability[k] = maxf(c.probability[k] / total, kSelectionFloor);
      floored += c.probability[k];
    }
  for (int k = 0; k < kGltfSlots; ++k) c.probability[k] = c.probability[k] / floored;
  return c;
}

// ComposedEvaluate, ComposedPdf, ComposedSample (composed.slang:335-445).
__device__ __forceinline__ V3 gltf_evaluate(const GltfClosure& c, V3 wo, V3 wi) {
  V3 value{0.0f, 0.0f, 0.0f};
  for (int k = 0; k < kGltfSlots; ++k) {
    if (!(c.present & (1u << k)) || !gltf_scatters_toward(k, wi)) continue;
    const LobeFrame f = lobe_frame(gltf_lobe_normal(c, k));
    value = value + c.weight[k] * gltf_lobe_value(c, k, to_lobe(f, wo), to_lobe(f, wi));
  }
  return value;
}
__device__ __forceinline__ float gltf_pdf(const GltfClosure& c, V3 wo, V3 wi) {
  float pdf = 0.0f;
  for (int k = 0; k < kGltfSlots; ++k) {
    if (!(c.present & (1u << k)) || !(c.probability[k] > 0.0f) || !gltf_scatters_toward(k, wi)) continue;
    const LobeFrame f = lobe_frame(gltf_lobe_normal(c, k));
    pdf += c.probability[k] * gltf_lobe_pdf(c, k, to_lobe(f, wo), to_lobe(f, wi));
  }
  return pdf;
}
__device__ __forceinline__ bool gltf_sample(const GltfClosure& c, V3 wo, float ux, float uy, float uz, BsdfSample& s) {
  s.wi = V3{0.0f, 0.0f, 1.0f};
  s.weight = V3{0.0f, 0.0f, 0.0f};
  s.pdf = 0.0f;
  s.delta = false;
  s.transmitted = false;
  int chosen = kGltfSlots;
  float start = 0.0f, end = 0.0f;
  for (int k = 0; k < kGltfSlots; ++k) {
    const float p = c.probability[k];
    if (!(p > 0.0f)) continue;
    chosen = k;
    start = end;
    end += p;
    if (uz < end) break;
  }
  if (chosen == kGltfSlots) return false;
  const LobeFrame f = lobe_frame(gltf_lobe_normal(c, chosen));
  V3 local;
  // the lobes here choose nothing with u.z: it only selected the lobe
  if (!gltf_lobe_sample(c, chosen, to_lobe(f, wo), ux, uy, local)) return false;
  s.wi = from_lobe(f, local);
  if (!gltf_scatters_toward(chosen, s.wi)) return false;
  s.transmitted = s.wi.z < 0.0f;
  s.pdf = gltf_pdf(c, wo, s.wi);
  if (!(s.pdf > 0.0f)) return false;
  s.weight = gltf_evaluate(c, wo, s.wi) / s.pdf;
  return finite3(s.weight);
}

// The 22 terms as a stored reservoir surface keeps them (gltf_pbr.slang:
// 159-214), the coat normal by its tangent components.
__device__ __forceinline__ void pack_gltf_terms(const GltfTerms& t, float l[22]) {
  l[0] = t.base_color.x, l[1] = t.base_color.y, l[2] = t.base_color.z, l[3] = t.alpha, l[4] = t.metallic;
  l[5] = t.e1, l[6] = t.share, l[7] = t.f0.x, l[8] = t.f0.y, l[9] = t.f0.z, l[10] = t.f90, l[11] = t.transmission;
  l[12] = t.coat_scale, l[13] = t.coat_dim, l[14] = t.coat_alpha, l[15] = t.coat_normal.x, l[16] = t.coat_normal.y;
  l[17] = t.sheen_color.x, l[18] = t.sheen_color.y, l[19] = t.sheen_color.z, l[20] = t.sheen_roughness,
  l[21] = t.sheen_dim;
}
__device__ __forceinline__ GltfTerms unpack_gltf_terms(const float l[22]) {
  GltfTerms t;
  t.base_color = V3{l[0], l[1], l[2]};
  t.alpha = l[3];
  t.metallic = l[4];
  t.e1 = l[5];
  t.share = l[6];
  t.f0 = V3{l[7], l[8], l[9]};
  t.f90 = l[10];
  t.transmission = l[11];
  t.coat_scale = l[12];
  t.coat_dim = l[13];
  t.coat_alpha = l[14];
  const float tx = l[15], ty = l[16];
  t.coat_normal = V3{tx, ty, sqrtf(maxf(1.0f - (tx * tx + ty * ty), 0.0f))};
  t.sheen_color = V3{l[17], l[18], l[19]};
  t.sheen_roughness = l[20];
  t.sheen_dim = l[21];
  return t;
}

}  // namespace lotus
