// One path step on the device: hit reconstruction, emission with MIS,
// Russian roulette, next event estimation and BSDF sampling. Behaviour follows
// backend/vulkan/shaders/common/transport.slang:164-247 and the scene access
// functions of common/scene.slang it calls (cited per function).
#pragma once

#include "bsdf.hpp"
#include "gltf.hpp"
#include "internal.hpp"

namespace lotus {

constexpr uint32_t kRouletteBounces = 3;  // transport.slang:21
constexpr float kMaxSurvival = 0.95f;     // transport.slang:22

// What every kernel of a frame reads besides the scene.
struct FrameConst {
  float clip_to_world[16];
  float world_to_clip[16];  // Vulkan clip space, for depth
  float ndc_from_pixel[4];
  float background[4];
  uint32_t width, height;
  uint32_t rect[4];
  uint32_t sample_index;
  uint32_t max_bounces;
  uint32_t light_sampling;
  uint32_t accumulated;  // samples already in the accumulation
  uint32_t output;
  // Radiance cache (cache_depth 0: off). A path whose pixel index plus
  // train_offset is a multiple of train_stride is a training path (index
  // (pixel + offset) / stride); every other path ends at bounce cache_depth
  // with a cache query. Positions map to [0, 1]^3 by (p - lo) * scale.
  uint32_t cache_depth;
  uint32_t train_stride, train_offset, train_capacity;
  float cache_lo[3], cache_scale[3];
  // ReSTIR DI (light_sampling kLightRestir): draws per first hit, proposal
  // (0 power table, 1 uniform, 2 mixture) and the records the first hit's
  // direct light reads (the last reuse pass's).
  uint32_t restir_candidates, restir_proposal;
  const float4* restir;
  // Temporal frames with reservoir reuse: receiver surfaces the initial pass
  // writes (8 float4 per pixel) and the camera's world -> view matrix.
  float4* restir_surf;
  float view[16];
  const uint64_t* instance_ids;  // stable ids (null: instance + 1, material slot + 1)
  const uint64_t* material_ids;
  // A replayed graph's frame variables in device memory (null: the fields
  // above): the sample index, the samples already accumulated and whether
  // this sample adds, written before each launch of the graph.
  const uint32_t* vars;
};

__device__ __forceinline__ uint32_t frame_sample_index(const FrameConst& F) {
  return F.vars ? F.vars[0] : F.sample_index;
}
__device__ __forceinline__ uint32_t frame_accumulated(const FrameConst& F) {
  return F.vars ? F.vars[1] : F.accumulated;
}

__device__ __forceinline__ uint64_t stable_instance_id(const FrameConst& F, uint32_t instance) {
  return F.instance_ids ? F.instance_ids[instance] : static_cast<uint64_t>(instance) + 1u;
}
__device__ __forceinline__ uint64_t stable_material_id(const FrameConst& F, const uint32_t* instance_material,
                                                       uint32_t instance) {
  return F.material_ids ? F.material_ids[instance] : static_cast<uint64_t>(instance_material[instance]) + 1u;
}

// The path state a kernel holds in registers.
struct PathRegs {
  V3 o, d;          // current ray
  V3 throughput;
  V3 radiance;
  uint32_t rng;
  float last_pdf;   // < 0: the last direction was delta or transmitted
  uint32_t depth;
};

// A hit's closure, whatever its model (bsdf.slang:23-90): UsdPreviewSurface's
// own or glTF PBR's composed one.
constexpr uint32_t kPreviewModel = 0u, kGltfModel = 1u;
struct SurfaceBsdf {
  uint32_t model;
  Closure preview;
  GltfClosure gltf;
};
__device__ __forceinline__ V3 bsdf_evaluate(const SurfaceBsdf& b, V3 wo, V3 wi) {
  return b.model == kGltfModel ? gltf_evaluate(b.gltf, wo, wi) : bsdf_evaluate(b.preview, wo, wi);
}
__device__ __forceinline__ float bsdf_pdf(const SurfaceBsdf& b, V3 wo, V3 wi) {
  return b.model == kGltfModel ? gltf_pdf(b.gltf, wo, wi) : bsdf_pdf(b.preview, wo, wi);
}
__device__ __forceinline__ bool bsdf_sample(const SurfaceBsdf& b, V3 wo, float ux, float uy, float uz,
                                            BsdfSample& s) {
  return b.model == kGltfModel ? gltf_sample(b.gltf, wo, ux, uy, uz, s) : bsdf_sample(b.preview, wo, ux, uy, uz, s);
}

struct Surface {
  V3 position, normal, shading_normal;
  float area;
  V3 base_color;
  float roughness, metallic;
  SurfaceBsdf bsdf;
  V3 emission;
};

// The surface summary of temporal products (scene.slang:1251-1258, 1396-1401):
// the albedo that demodulates its light (UsdPreviewSurface's transmitted
// share tinted by the transmission colour) and the principal lobe's
// roughness and share of the reflection.
__device__ __forceinline__ V3 surface_albedo(const Surface& sf) {
  if (sf.bsdf.model == kGltfModel) return sf.base_color;
  const Closure& b = sf.bsdf.preview;
  return b.transmission > 0.0f ? sf.base_color + (b.transmission_color - sf.base_color) * b.transmission
                               : sf.base_color;
}
__device__ __forceinline__ float surface_specular_share(const Surface& sf) {
  if (sf.bsdf.model == kGltfModel)
    return sf.metallic + max3(sf.bsdf.gltf.t.f0) * (1.0f - sf.metallic);
  return sf.metallic + 0.04f * (1.0f - sf.metallic);
}

__device__ __forceinline__ float word_float(const uint32_t* w, int i) { return __uint_as_float(w[i]); }

// scene.slang:273-283.
__device__ __forceinline__ Ray camera_ray(const float* clip_to_world, float nx, float ny) {
  float nh[4], fh[4];
  mul_point(clip_to_world, nx, ny, 0.0f, 1.0f, nh);
  mul_point(clip_to_world, nx, ny, 1.0f, 1.0f, fh);
  Ray r;
  r.o = V3{nh[0] / nh[3], nh[1] / nh[3], nh[2] / nh[3]};
  r.d = normalize3(V3{fh[0], fh[1], fh[2]} * nh[3] - V3{nh[0], nh[1], nh[2]} * fh[3]);
  r.tmin = 0.0f;
  r.tmax = fh[3] == 0.0f ? kFloatMax : length3(V3{fh[0] / fh[3], fh[1] / fh[3], fh[2] / fh[3]} - r.o);
  return r;
}

// transport.slang:53-69 (StartPath): box-filtered film position.
__device__ __forceinline__ Ray start_path(const FrameConst& f, uint32_t x, uint32_t y, uint32_t sample,
                                          uint32_t& rng) {
  rng = seed_random(y * f.width + x, sample);
  const float ux = next_random(rng), uy = next_random(rng);
  const float fx = (static_cast<float>(x) + 0.5f) + (ux - 0.5f);
  const float fy = (static_cast<float>(y) + 0.5f) + (uy - 0.5f);
  return camera_ray(f.clip_to_world, fx * f.ndc_from_pixel[0] + f.ndc_from_pixel[2],
                    fy * f.ndc_from_pixel[1] + f.ndc_from_pixel[3]);
}

// ---------------------------------------------------- material textures
// Texture-input slots (scene.slang:59-67), lookup modes (gpu_scene.hpp:56-59)
// and wrap modes (TextureWrap, material.hpp:43-50).
constexpr uint32_t kOpacitySlot = 0, kNormalSlot = 1, kBaseColorSlot = 2, kRoughnessSlot = 3, kMetallicSlot = 4, kEmissionSlot = 5,
                   kSpecularColorSlot = 6;
constexpr uint32_t kPreviewLookupSlots = (1u << kNormalSlot) | (1u << kBaseColorSlot) | (1u << kRoughnessSlot) |
                                         (1u << kMetallicSlot) | (1u << kEmissionSlot) | (1u << kSpecularColorSlot);
constexpr uint32_t kNoCorner = 0xFFFFFFFFu;
constexpr uint32_t kConstantColor = 0x80000000u;
constexpr uint32_t kFallbackLookup = 2;
constexpr uint32_t kWrapBlack = 0, kWrapRepeat = 2, kWrapMirror = 3;

// A texel index under a wrap mode; `black` is set outside a black border.
__device__ __forceinline__ int wrap_texel(int i, int n, uint32_t mode, bool& black) {
  if (mode == kWrapRepeat) {
    const int r = i % n;
    return r < 0 ? r + n : r;
  }
  if (mode == kWrapMirror) {
    int r = i % (2 * n);
    if (r < 0) r += 2 * n;
    return r < n ? r : 2 * n - 1 - r;
  }
  if (i < 0 || i >= n) {
    if (mode == kWrapBlack) black = true;
    return i < 0 ? 0 : n - 1;
  }
  return i;
}

// A texel coordinate kept within +-2^24, where floats are still integers
// and the corner indices below fit an int; NaN (coordinates that overflowed
// on the way) reads texel 0. Coordinates of real lookups lie far inside.
__device__ __forceinline__ float lookup_coordinate(float v) {
  return v == v ? minf(maxf(v, -16777216.0f), 16777216.0f) : 0.0f;
}

// scene.slang:318-329 (Lookup): the bilinear texel at (s', 1 - t') of the
// binding's coordinate transform, times its scale plus its bias; or its
// fallback. Binding words: texture, sampler (wrap_s * 4 + wrap_t), channel,
// mode, scale[4], bias[4], fallback[4], st_s[4], st_t[4].
__device__ __forceinline__ float4 texture_lookup(const SceneView& S, const uint32_t* b, float s, float t) {
  const float* f = reinterpret_cast<const float*>(b);
  if (b[3] == kFallbackLookup) return make_float4(f[12], f[13], f[14], f[15]);
  const float ss = (f[16] * s + f[17] * t) + f[18];
  const float tt = 1.0f - ((f[20] * s + f[21] * t) + f[22]);
  const uint4 tex = S.textures[b[0]];
  const int W = static_cast<int>(tex.y), H = static_cast<int>(tex.z);
  const float x = lookup_coordinate(ss * static_cast<float>(W) - 0.5f),
              y = lookup_coordinate(tt * static_cast<float>(H) - 0.5f);
  const float x0 = floorf(x), y0 = floorf(y);
  const float fx = x - x0, fy = y - y0;
  const uint32_t ws = b[1] >> 2, wt = b[1] & 3u;
  auto texel = [&](int i, int j) {
    bool black = false;
    const int xi = wrap_texel(i, W, ws, black), yj = wrap_texel(j, H, wt, black);
    return black ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : S.texels[tex.x + static_cast<uint32_t>(yj * W + xi)];
  };
  const int i0 = static_cast<int>(x0), j0 = static_cast<int>(y0);
  const float4 t00 = texel(i0, j0), t10 = texel(i0 + 1, j0), t01 = texel(i0, j0 + 1), t11 = texel(i0 + 1, j0 + 1);
  auto mix = [&](float a, float bb, float c, float d) {
    return ((1.0f - fx) * a + fx * bb) * (1.0f - fy) + ((1.0f - fx) * c + fx * d) * fy;
  };
  return make_float4(mix(t00.x, t10.x, t01.x, t11.x) * f[4] + f[8], mix(t00.y, t10.y, t01.y, t11.y) * f[5] + f[9],
                     mix(t00.z, t10.z, t01.z, t11.z) * f[6] + f[10], mix(t00.w, t10.w, t01.w, t11.w) * f[7] + f[11]);
}

// scene.slang:345-362 (MaterialInput): slot `slot`'s binding, or null for a
// constant input.
__device__ __forceinline__ const uint32_t* material_input(const SceneView& S, const uint32_t* m, uint32_t slot) {
  const uint32_t bit = 1u << slot;
  if ((m[2] & bit) == 0) return nullptr;
  return S.bindings + 24ull * (m[1] + static_cast<uint32_t>(__popc(m[2] & (bit - 1u))));
}

// scene.slang:364-374 (HitTexcoords): the instance's corners interpolated
// with the barycentrics, or the material's fallback coordinates.
__device__ __forceinline__ void hit_texcoords(const SceneView& S, const uint32_t* m, uint32_t instance,
                                              uint32_t primitive, float u, float v, float& s, float& t) {
  const uint32_t base = S.instance_texcoords[instance];
  if (base == 0xFFFFFFFFu) {
    s = word_float(m, 4);
    t = word_float(m, 5);
    return;
  }
  const float2 t0 = S.texcoords[base + 3ull * primitive], t1 = S.texcoords[base + 3ull * primitive + 1],
               t2 = S.texcoords[base + 3ull * primitive + 2];
  s = (t0.x + u * (t1.x - t0.x)) + v * (t2.x - t0.x);
  t = (t0.y + u * (t1.y - t0.y)) + v * (t2.y - t0.y);
}

__device__ __forceinline__ float channel_of(float4 v, uint32_t c) { return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w; }

// scene.slang:393-406 (Coverage): the probability that a ray accepts a
// candidate of `instance` at barycentrics (u, v): its material's opacity,
// constant or looked up (clamped to [0, 1], 0 where the lookup is not
// finite), or 0 or 1 against a positive threshold.
__device__ __forceinline__ float surface_coverage(const SceneView& S, uint32_t instance, uint32_t primitive, float u,
                                                  float v) {
  const uint32_t* m = S.materials + 36ull * S.instance_material[instance];
  float opacity = word_float(m, 6);
  if (const uint32_t* b = material_input(S, m, kOpacitySlot)) {
    float s, t;
    hit_texcoords(S, m, instance, primitive, u, v, s, t);
    const float o = channel_of(texture_lookup(S, b, s, t), b[2]);
    opacity = isfinite(o) ? saturate(o) : 0.0f;
  }
  const float threshold = word_float(m, 7);
  if (threshold > 0.0f) return opacity >= threshold ? 1.0f : 0.0f;
  return opacity;
}

// The coverage policy of walks over a scene with covered triangles.
struct SceneCoverage {
  static constexpr bool kEnabled = true;
  const SceneView* S;
  __device__ __forceinline__ float operator()(uint32_t instance, uint32_t primitive, float u, float v) const {
    return surface_coverage(*S, instance, primitive, u, v);
  }
};

// scene.slang:572-577 (PreviewEmission) at texture coordinates (s, t).
__device__ __forceinline__ V3 preview_emission(const SceneView& S, const uint32_t* m, float s, float t) {
  const uint32_t* b = material_input(S, m, kEmissionSlot);
  if (!b) return V3{word_float(m, 16), word_float(m, 17), word_float(m, 18)};
  const float4 e = texture_lookup(S, b, s, t);
  return V3{maxf(e.x, 0.0f), maxf(e.y, 0.0f), maxf(e.z, 0.0f)};
}

// scene.slang:467-488 (ShadingNormal): the corner normals at `corner`
// interpolated at the hit, carried to world space by the cofactor matrix of
// the instance's linear part, turned to the geometric normal's side; the
// geometric normal where the result vanishes or faces away from the ray.
__device__ __forceinline__ V3 authored_normal(const SceneView& S, uint32_t corner, float u, float v, uint32_t instance,
                                              V3 geometric, V3 dir) {
  const float* n = S.normals + 3ull * corner;
  const V3 n0{n[0], n[1], n[2]}, n1{n[3], n[4], n[5]}, n2{n[6], n[7], n[8]};
  const V3 object = (n0 + u * (n1 - n0)) + v * (n2 - n0);
  const float* x = S.instance_transform + 16ull * instance;
  float a[4], b[4], c[4];
  mul_point(x, 1.0f, 0.0f, 0.0f, 0.0f, a);
  mul_point(x, 0.0f, 1.0f, 0.0f, 0.0f, b);
  mul_point(x, 0.0f, 0.0f, 1.0f, 0.0f, c);
  const V3 c0{a[0], a[1], a[2]}, c1{b[0], b[1], b[2]}, c2{c[0], c[1], c[2]};
  const V3 world = (cross3(c1, c2) * object.x + cross3(c2, c0) * object.y) + cross3(c0, c1) * object.z;
  const float length2 = dot3(world, world);
  if (!(length2 > 0.0f) || isinf(length2)) return geometric;
  V3 shading = world * rsqrt_rn(length2);
  if (dot3(shading, geometric) < 0.0f) shading = -shading;
  return dot3(shading, dir) < 0.0f ? shading : geometric;
}

// scene.slang:376-385 (HitColor): the instance's colour set at the hit.
__device__ __forceinline__ V3 hit_color(const SceneView& S, uint32_t set, uint32_t primitive, float u, float v) {
  if (set & kConstantColor) {
    const float* c = S.colors + 3ull * (set & ~kConstantColor);
    return V3{c[0], c[1], c[2]};
  }
  const float* c = S.colors + 3ull * (uint64_t(set) + 3ull * primitive);
  const V3 c0{c[0], c[1], c[2]}, c1{c[3], c[4], c[5]}, c2{c[6], c[7], c[8]};
  return (c0 + u * (c1 - c0)) + v * (c2 - c0);
}

// scene.slang:490-537 (TangentFrame): the shading frame about `normal` whose
// tangent follows the texture coordinates' s direction (the normal map's
// transform applied to their differentials), or the normal-only frame
// (MakeShadingFrame, sampling.slang:45-56) without usable coordinates.
__device__ __forceinline__ void tangent_frame(const SceneView& S, const uint32_t* input, uint32_t instance,
                                              uint32_t primitive, V3 edge1, V3 edge2, V3 normal, V3& ft, V3& fb) {
  {
    const V3 n = normal;
    const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    ft = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    fb = V3{b, sign + n.y * n.y * a, -n.y};
  }
  const uint32_t first = S.instance_texcoords ? S.instance_texcoords[instance] : kNoCorner;
  if (first == kNoCorner) return;
  const float2* t = S.texcoords + (uint64_t(first) + 3ull * primitive);
  const float2 st0 = t[0];
  float u1x = t[1].x - st0.x, u1y = t[1].y - st0.y;
  float u2x = t[2].x - st0.x, u2y = t[2].y - st0.y;
  if (input) {
    const float* f = reinterpret_cast<const float*>(input);
    const float a1 = f[16] * u1x + f[17] * u1y, b1 = f[20] * u1x + f[21] * u1y;
    const float a2 = f[16] * u2x + f[17] * u2y, b2 = f[20] * u2x + f[21] * u2y;
    u1x = a1;
    u1y = b1;
    u2x = a2;
    u2y = b2;
  }
  const float uv_scale = maxf(maxf(fabsf(u1x), fabsf(u1y)), maxf(fabsf(u2x), fabsf(u2y)));
  const float edge_scale = maxf(maxf(maxf(fabsf(edge1.x), fabsf(edge1.y)), fabsf(edge1.z)),
                                maxf(maxf(fabsf(edge2.x), fabsf(edge2.y)), fabsf(edge2.z)));
  if (!(uv_scale > 0.0f) || isinf(uv_scale) || !(edge_scale > 0.0f) || isinf(edge_scale)) return;
  u1x = u1x / uv_scale;
  u1y = u1y / uv_scale;
  u2x = u2x / uv_scale;
  u2y = u2y / uv_scale;
  edge1 = edge1 / edge_scale;
  edge2 = edge2 / edge_scale;
  const float determinant = u1x * u2y - u1y * u2x;
  if (!(fabsf(determinant) > 1.0e-8f)) return;
  const float orientation = determinant < 0.0f ? -1.0f : 1.0f;
  V3 tangent = orientation * (edge1 * u2y - edge2 * u1y);
  const V3 bitangent = orientation * (edge2 * u1x - edge1 * u2x);
  tangent = tangent - normal * dot3(normal, tangent);
  const float length2 = dot3(tangent, tangent);
  if (!(length2 > 1.0e-16f) || isinf(length2)) return;
  ft = tangent * rsqrt_rn(length2);
  fb = cross3(normal, ft);
  if (dot3(fb, bitangent) < 0.0f) fb = -fb;
}

// MakeShadingFrame (sampling.slang:45-56, Duff et al. 2017): tangent and
// bitangent of unit n.
__device__ __forceinline__ void shading_frame(V3 n, V3& ft, V3& fb) {
  const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  ft = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
  fb = V3{b, sign + n.y * n.y * a, -n.y};
}

// scene.slang:541-553 (MappedNormal) of tangent-space normal `tn` in the
// tangent frame about `base` (TangentFrame with the map's binding `input`,
// null for a constant normal): the shading normal, or `base` where the mapped
// one faces away from the ray.
template <bool kExtras>
__device__ __forceinline__ V3 map_normal(const SceneView& S, V3 tn, const uint32_t* input, uint32_t instance,
                                         uint32_t primitive, V3 e1, V3 e2, V3 base, V3 geometric, V3 dir) {
  V3 bt, bb;
  if (kExtras) {
    tangent_frame(S, input, instance, primitive, e1, e2, base, bt, bb);
  } else {
    shading_frame(base, bt, bb);
  }
  const float mag = maxf(maxf(fabsf(tn.x), fabsf(tn.y)), fabsf(tn.z));
  if (!(mag > 0.0f) || !finite3(tn)) return base;
  const V3 q = tn / mag;
  V3 sh = normalize3(((q.x * bt) + (q.y * bb)) + (q.z * base));
  if (dot3(sh, geometric) < 0.0f) sh = -sh;
  return dot3(sh, dir) < 0.0f ? sh : base;
}

// A scalar glTF input (scene.slang:1262-1274): its lookup's channel, clamped,
// where the slot has one, else the constant.
__device__ __forceinline__ float gltf_scalar(const SceneView& S, const uint32_t* m, uint32_t slot, float s, float t,
                                             float constant) {
  const uint32_t* b = material_input(S, m, slot);
  return b ? saturate(channel_of(texture_lookup(S, b, s, t), b[2])) : constant;
}

__device__ __forceinline__ V3 clamp_normal(float4 c) {
  return V3{minf(maxf(c.x, -1.0f), 1.0f), minf(maxf(c.y, -1.0f), 1.0f), minf(maxf(c.z, -1.0f), 1.0f)};
}

// scene.slang:1276-1403 (EvaluateGltfPbr) into `sf`, whose geometry and
// shading normal are set: the inputs (material words 12-33,
// gpu_scene.cpp:2050-2071) and their lookups (slots: normal 1, base colour
// 2, metallic 3, roughness 4, emissive 5, specular 6, specular colour 7,
// transmission 8, clearcoat 9, its roughness 10, its normal 11, sheen colour
// 12, sheen roughness 13), the base's and the coat's normal maps, the
// closure's outgoing direction (the shading frame's x-z plane without a
// coat), its terms and its closure.
__device__ __forceinline__ void gltf_surface(const SceneView& S, const uint32_t* m, uint32_t instance,
                                             uint32_t primitive, float u, float v, V3 e1, V3 e2, V3 dir,
                                             Surface& sf) {
  V3 base = V3{word_float(m, 12), word_float(m, 13), word_float(m, 14)};
  float roughness = word_float(m, 15), metallic = word_float(m, 19);
  const float strength = word_float(m, 20);
  V3 emission = V3{word_float(m, 16), word_float(m, 17), word_float(m, 18)} * strength;
  V3 tn = V3{word_float(m, 8), word_float(m, 9), word_float(m, 10)};
  float specular = word_float(m, 22), transmission = word_float(m, 23), clearcoat = word_float(m, 27);
  float sheen_roughness = word_float(m, 31), clearcoat_roughness = word_float(m, 32);
  V3 specular_color = V3{word_float(m, 24), word_float(m, 25), word_float(m, 26)};
  V3 sheen_color = V3{word_float(m, 28), word_float(m, 29), word_float(m, 30)};
  const V3 mesh_normal = sf.shading_normal;
  V3 coat_tn{0.0f, 0.0f, 1.0f};
  const uint32_t* normal_input = nullptr;
  const uint32_t* coat_input = nullptr;
  float s = 0.0f, tc = 0.0f;
  if (m[2] & 0x3FFEu) hit_texcoords(S, m, instance, primitive, u, v, s, tc);
  if (m[2] & 0x3Eu) {
    if (const uint32_t* b = material_input(S, m, 2)) {
      const float4 c = texture_lookup(S, b, s, tc);
      base = V3{saturate(c.x), saturate(c.y), saturate(c.z)};
    }
    if (const uint32_t* b = material_input(S, m, 4)) roughness = saturate(channel_of(texture_lookup(S, b, s, tc), b[2]));
    if (const uint32_t* b = material_input(S, m, 3)) metallic = saturate(channel_of(texture_lookup(S, b, s, tc), b[2]));
    if (const uint32_t* b = material_input(S, m, 5)) {
      const float4 c = texture_lookup(S, b, s, tc);
      emission = V3{maxf(c.x, 0.0f), maxf(c.y, 0.0f), maxf(c.z, 0.0f)} * strength;
    }
    if (const uint32_t* b = material_input(S, m, 1)) {
      tn = clamp_normal(texture_lookup(S, b, s, tc));
      normal_input = b;
    }
  }
  if (m[2] & 0x3FC0u) {
    specular = gltf_scalar(S, m, 6, s, tc, specular);
    transmission = gltf_scalar(S, m, 8, s, tc, transmission);
    clearcoat = gltf_scalar(S, m, 9, s, tc, clearcoat);
    clearcoat_roughness = gltf_scalar(S, m, 10, s, tc, clearcoat_roughness);
    sheen_roughness = gltf_scalar(S, m, 13, s, tc, sheen_roughness);
    if (const uint32_t* b = material_input(S, m, 7)) {
      const float4 c = texture_lookup(S, b, s, tc);
      specular_color = V3{maxf(c.x, 0.0f), maxf(c.y, 0.0f), maxf(c.z, 0.0f)};
    }
    if (const uint32_t* b = material_input(S, m, 12)) {
      const float4 c = texture_lookup(S, b, s, tc);
      sheen_color = V3{saturate(c.x), saturate(c.y), saturate(c.z)};
    }
    if (const uint32_t* b = material_input(S, m, 11)) {
      coat_tn = clamp_normal(texture_lookup(S, b, s, tc));
      coat_input = b;
    }
  }
  if (!(tn.x == 0.0f && tn.y == 0.0f && tn.z == 1.0f))
    sf.shading_normal = map_normal<true>(S, tn, normal_input, instance, primitive, e1, e2, sf.shading_normal,
                                         sf.normal, dir);
  const float cos_o = dot3(sf.shading_normal, -dir);
  V3 wo{sqrtf(maxf(1.0f - cos_o * cos_o, 0.0f)), 0.0f, cos_o};
  GltfInputs in;
  in.base_color = base;
  in.roughness = roughness;
  in.metallic = metallic;
  in.dielectric_f0 = word_float(m, 21);
  in.specular = specular;
  in.specular_color = specular_color;
  in.transmission = transmission;
  in.clearcoat = clearcoat;
  in.clearcoat_roughness = clearcoat_roughness;
  in.clearcoat_normal = V3{0.0f, 0.0f, 1.0f};
  in.sheen_color = sheen_color;
  in.sheen_roughness = sheen_roughness;
  if (clearcoat > 0.0f) {
    V3 coat_normal = mesh_normal;
    if (!(coat_tn.x == 0.0f && coat_tn.y == 0.0f && coat_tn.z == 1.0f))
      coat_normal = map_normal<true>(S, coat_tn, coat_input, instance, primitive, e1, e2, mesh_normal, sf.normal, dir);
    V3 ft, fb;
    shading_frame(sf.shading_normal, ft, fb);
    const V3 md = -dir;
    wo.x = dot3(md, ft);
    wo.y = dot3(md, fb);
    in.clearcoat_normal = normalize3(V3{dot3(coat_normal, ft), dot3(coat_normal, fb), dot3(coat_normal, sf.shading_normal)});
  }
  sf.bsdf.model = kGltfModel;
  sf.bsdf.gltf = make_gltf_closure(make_gltf_terms(S.albedo_tables, in, wo), wo);
  sf.emission = emission;
  sf.base_color = base;
  sf.roughness = roughness;
  sf.metallic = metallic;
}

// scene.slang:1408-1449 with EvaluatePreview (:1184-1260): authored shading
// normals, vertex colours, texture lookups (with kExtras) and MappedNormal
// (:541-553) for the tangent-space normal, constant or mapped; glTF PBR
// materials (with kExtras) through gltf_surface.
// kGltf: glTF PBR's closure compiled in (a variant of its own: its closure
// needs far more registers than UsdPreviewSurface's)
template <bool kExtras = true, bool kGltf = kExtras>  // normals, colours and lookups compiled in
__device__ __forceinline__ Surface reconstruct_hit(const SceneView& S, uint32_t tri, uint32_t instance,
                                                   float u, float v, V3 dir) {
  const Tri t = S.tris[tri];
  const V3 p0 = xyz(t.a), p1 = xyz(t.b), p2 = xyz(t.c);
  Surface sf;
  sf.position = (p0 + u * (p1 - p0)) + v * (p2 - p0);
  const V3 e = cross3(p1 - p0, p2 - p0);
  const V3 normal = normalize3(e);
  sf.area = 0.5f * length3(e);
  sf.normal = dot3(normal, dir) > 0.0f ? -normal : normal;
  sf.shading_normal = sf.normal;
  const uint32_t primitive = __float_as_uint(t.c.w) & ~kCoveredBit;
  if (kExtras && S.instance_normals) {
    const uint32_t first = S.instance_normals[instance];
    if (first != kNoCorner)
      sf.shading_normal = authored_normal(S, first + 3u * primitive, u, v, instance, sf.normal, dir);
  }
  const uint32_t* m = S.materials + 36ull * S.instance_material[instance];
  sf.bsdf.model = kPreviewModel;
  if (kGltf && (m[0] & 0xFFu) == kGltfModel) {
    gltf_surface(S, m, instance, primitive, u, v, p1 - p0, p2 - p0, dir, sf);
    return sf;
  }
  const uint32_t flags = m[0] >> 8;
  V3 tn = V3{word_float(m, 8), word_float(m, 9), word_float(m, 10)};
  sf.base_color = V3{word_float(m, 12), word_float(m, 13), word_float(m, 14)};
  if (kExtras && S.instance_colors) {
    const uint32_t set = S.instance_colors[instance];
    if (set != kNoCorner) {
      const V3 c = hit_color(S, set, primitive, u, v);
      sf.base_color = V3{saturate(c.x), saturate(c.y), saturate(c.z)};
    }
  }
  sf.roughness = word_float(m, 15);
  sf.emission = V3{word_float(m, 16), word_float(m, 17), word_float(m, 18)};
  sf.metallic = word_float(m, 19);
  V3 specular = V3{word_float(m, 20), word_float(m, 21), word_float(m, 22)};
  const uint32_t* normal_input = nullptr;
  if (kExtras && (m[2] & kPreviewLookupSlots)) {
    // a lookup replaces its input's constant, clamped into its range
    float s, tc;
    hit_texcoords(S, m, instance, primitive, u, v, s, tc);
    if (const uint32_t* b = material_input(S, m, kBaseColorSlot)) {
      const float4 c = texture_lookup(S, b, s, tc);
      sf.base_color = V3{saturate(c.x), saturate(c.y), saturate(c.z)};
    }
    if (const uint32_t* b = material_input(S, m, kRoughnessSlot))
      sf.roughness = saturate(channel_of(texture_lookup(S, b, s, tc), b[2]));
    if (const uint32_t* b = material_input(S, m, kMetallicSlot))
      sf.metallic = saturate(channel_of(texture_lookup(S, b, s, tc), b[2]));
    sf.emission = preview_emission(S, m, s, tc);
    if (const uint32_t* b = material_input(S, m, kNormalSlot)) {
      const float4 c = texture_lookup(S, b, s, tc);
      tn = V3{minf(maxf(c.x, -1.0f), 1.0f), minf(maxf(c.y, -1.0f), 1.0f), minf(maxf(c.z, -1.0f), 1.0f)};
      normal_input = b;
    }
    if (const uint32_t* b = material_input(S, m, kSpecularColorSlot)) {
      const float4 c = texture_lookup(S, b, s, tc);
      const V3 value{c.x, c.y, c.z};
      if (finite3(value)) specular = V3{saturate(value.x), saturate(value.y), saturate(value.z)};
    }
  }
  if (!(tn.x == 0.0f && tn.y == 0.0f && tn.z == 1.0f))
    sf.shading_normal = map_normal<kExtras>(S, tn, normal_input, instance, primitive, p1 - p0, p2 - p0,
                                            sf.shading_normal, sf.normal, dir);
  const float ior = word_float(m, 23);
  const V3 tcolor = V3{word_float(m, 24), word_float(m, 25), word_float(m, 26)};
  const float transmission = word_float(m, 27);
  bool entering = true;
  if (transmission > 0.0f) {
    const float* x = S.instance_transform + 16ull * instance;
    float c0[4], c1[4], c2[4];
    mul_point(x, 1.0f, 0.0f, 0.0f, 0.0f, c0);
    mul_point(x, 0.0f, 1.0f, 0.0f, 0.0f, c1);
    mul_point(x, 0.0f, 0.0f, 1.0f, 0.0f, c2);
    const V3 outward = dot3(V3{c0[0], c0[1], c0[2]}, cross3(V3{c1[0], c1[1], c1[2]}, V3{c2[0], c2[1], c2[2]})) < 0.0f
                           ? -normal
                           : normal;
    entering = dot3(outward, dir) < 0.0f;
  }
  sf.bsdf.preview = make_preview_closure(sf.base_color, sf.roughness, sf.metallic, ior, (flags & 1u) != 0, specular,
                                 transmission, tcolor, (flags & 2u) != 0, entering);
  return sf;
}

// sampling.slang:67-76.
__device__ __forceinline__ float power_heuristic(float chosen, float other) {
  if (!(chosen > 0.0f)) return 0.0f;
  if (isinf(chosen)) return isinf(other) ? 0.5f : 1.0f;
  const float ratio = other / chosen;
  return 1.0f / (1.0f + ratio * ratio);
}

struct LightSample {
  V3 direction;
  float distance;
  V3 radiance;
  float pdf;
  bool delta;  // a parallel light: only next event estimation finds it
};

// scene.slang:632-677 (SampleTriangle): a point uniformly by area on light
// `light` of the table, (uz, uw) its unit-square coordinates; kCov compiles
// in the coverage of covered emitters.
template <bool kExtras = true, bool kCov = kExtras>
__device__ __forceinline__ LightSample sample_triangle_light(const SceneView& S, uint32_t light, float uz, float uw,
                                                             V3 origin) {
  LightSample none{V3{0.0f, 0.0f, 1.0f}, 0.0f, V3{0.0f, 0.0f, 0.0f}, 0.0f};
  const LightRecord chosen = S.lights[light];
  const uint32_t prim = light - S.light_offsets[chosen.instance];
  const float root = sqrtf(uz);
  const float bu = root * (1.0f - uw), bv = root * uw;
  const uint32_t g = S.instance_base[chosen.instance] + prim;
  const Tri t = S.tris[S.tri_of_global[g]];
  const V3 p0 = xyz(t.a), p1 = xyz(t.b), p2 = xyz(t.c);
  const uint32_t* m = S.materials + 36ull * S.instance_material[chosen.instance];
  const V3 position = (p0 + bu * (p1 - p0)) + bv * (p2 - p0);
  const V3 e = cross3(p1 - p0, p2 - p0);
  const V3 normal = normalize3(e);
  const float area = 0.5f * length3(e);
  // Emission (scene.slang:591-600): glTF PBR's emissive times its strength
  const bool gltf = kExtras && (m[0] & 0xFFu) == kGltfModel;
  V3 emission = V3{word_float(m, 16), word_float(m, 17), word_float(m, 18)};
  if (gltf) emission = emission * word_float(m, 20);
  if (kExtras && (m[2] & (1u << kEmissionSlot))) {
    float s, tc;
    hit_texcoords(S, m, chosen.instance, prim, bu, bv, s, tc);
    emission = preview_emission(S, m, s, tc);
    if (gltf) emission = emission * word_float(m, 20);
  }
  // an emitter's radiance times its coverage, the probability that a ray
  // reaching it stops there (scene.slang:608-611, 656-657)
  const V3 radiance =
      emission * ((kCov && (__float_as_uint(t.c.w) & kCoveredBit)) ? surface_coverage(S, chosen.instance, prim, bu, bv)
                                                                    : 1.0f);
  const V3 toward = position - origin;
  const float d2 = dot3(toward, toward);
  if (!(area > 0.0f) || !(d2 > 0.0f) || max3(radiance) <= 0.0f) return none;
  LightSample ls;
  ls.delta = false;
  ls.direction = toward * rsqrt_rn(d2);
  const float cosine = fabsf(dot3(normal, ls.direction));
  if (!(cosine > 0.0f)) return none;
  const V3 end = offset_ray_origin(position, dot3(normal, ls.direction) > 0.0f ? -normal : normal);
  ls.distance = maxf(0.0f, dot3(end - origin, ls.direction));
  ls.radiance = radiance;
  ls.pdf = chosen.pdf / area * d2 / cosine;
  return ls;
}

__device__ __forceinline__ V3 to_world(V3 v, V3 ft, V3 fb, V3 n) { return ((v.x * ft) + (v.y * fb)) + (v.z * n); }

// scene.slang:682-689: the solid-angle density of a sphere's cone, given
// sin^2 of its half angle (1 - cos without cancellation for small spheres).
__device__ __forceinline__ float sphere_cone_pdf(float sin2_max) {
  const float one_minus_cos = sin2_max < 0.00068523f ? 0.5f * sin2_max : 1.0f - sqrtf(maxf(0.0f, 1.0f - sin2_max));
  return 1.0f / ((2.0f * kPi) * one_minus_cos);
}

// sampling.slang:89-104 (SampleConcentricDisk; Shirley and Chiu), the angle
// as a fraction of a turn.
__device__ __forceinline__ void concentric_disk(float ux, float uy, float& x, float& y) {
  const float ox = 2.0f * ux - 1.0f, oy = 2.0f * uy - 1.0f;
  if (ox == 0.0f && oy == 0.0f) {
    x = y = 0.0f;
    return;
  }
  float radius, turns;
  if (fabsf(ox) > fabsf(oy)) {
    radius = ox;
    turns = 0.125f * (oy / ox);
  } else {
    radius = oy;
    turns = 0.25f - 0.125f * (ox / oy);
  }
  float sn, cs;
  sincos_turns(turns, &sn, &cs);
  x = radius * cs;
  y = radius * sn;
}

// scene.slang:693-760 (SampleAnalytic): a direction towards analytic light
// `L` chosen with probability `selection`: by solid angle over a sphere's
// cone or a distant light's, by area on a rect or disk; a parallel light's
// one direction.
__device__ __forceinline__ LightSample sample_analytic(const AnalyticRecord& L, float selection, float ux, float uy,
                                                       V3 origin) {
  const LightSample none{V3{0.0f, 0.0f, 1.0f}, 0.0f, V3{0.0f, 0.0f, 0.0f}, 0.0f, false};
  LightSample ls = none;
  ls.radiance = V3{L.radiance[0], L.radiance[1], L.radiance[2]};
  const uint32_t shape = L.shape;
  if (shape == kSphereLight) {
    // uniform over the cone, then the cone direction's point on the near
    // side (Pharr et al., Physically Based Rendering, 4th ed., 6.2.4)
    const V3 centre = V3{L.position[0], L.position[1], L.position[2]};
    const float radius = L.position[3];
    const V3 to_centre = centre - origin;
    const float d2 = dot3(to_centre, to_centre);
    if (!(d2 > radius * radius)) return none;
    const float sin2_max = radius * radius / d2;
    const float sin_max = sqrtf(sin2_max);
    const float cos_max = sqrtf(maxf(0.0f, 1.0f - sin2_max));
    float cos_theta = (cos_max - 1.0f) * ux + 1.0f;
    float sin2_theta = 1.0f - cos_theta * cos_theta;
    if (sin2_max < 0.00068523f) {
      sin2_theta = sin2_max * ux;
      cos_theta = sqrtf(1.0f - sin2_theta);
    }
    const float cos_alpha = sin2_theta / sin_max + cos_theta * sqrtf(maxf(0.0f, 1.0f - sin2_theta / sin2_max));
    const float sin_alpha = sqrtf(maxf(0.0f, 1.0f - cos_alpha * cos_alpha));
    float sp, cp;
    sincos_turns(uy, &sp, &cp);
    const V3 n = normalize3(to_centre);
    V3 ft, fb;
    shading_frame(n, ft, fb);
    const V3 outward = to_world(-V3{sin_alpha * cp, sin_alpha * sp, cos_alpha}, ft, fb, n);
    const V3 toward = (centre + radius * outward) - origin;
    ls.distance = length3(toward);
    ls.direction = toward / ls.distance;
    ls.pdf = selection * sphere_cone_pdf(sin2_max);
  } else if (shape == kRectLight || shape == kDiskLight) {
    float a, b;
    if (shape == kRectLight) {
      a = 2.0f * ux - 1.0f;
      b = 2.0f * uy - 1.0f;
    } else {
      concentric_disk(ux, uy, a, b);
    }
    const V3 toward = ((V3{L.position[0], L.position[1], L.position[2]} + a * V3{L.axis_u[0], L.axis_u[1], L.axis_u[2]}) +
                       b * V3{L.axis_v[0], L.axis_v[1], L.axis_v[2]}) -
                      origin;
    const float d2 = dot3(toward, toward);
    if (!(d2 > 0.0f)) return none;
    ls.distance = sqrtf(d2);
    ls.direction = toward / ls.distance;
    const float cosine = -dot3(V3{L.normal[0], L.normal[1], L.normal[2]}, ls.direction);
    if (!(cosine > 0.0f)) return none;  // light leaves only its emitting side
    ls.pdf = selection * d2 / (cosine * L.normal[3]);
  } else if (shape == kDistantLight) {
    // SampleUniformCone (sampling.slang:109-116) about the light's direction
    const V3 n = V3{L.axis_u[0], L.axis_u[1], L.axis_u[2]};
    const float one_minus_cos = ux * L.axis_u[3];
    const float sin_theta = sqrtf(maxf(0.0f, one_minus_cos * (2.0f - one_minus_cos)));
    float sp, cp;
    sincos_turns(uy, &sp, &cp);
    V3 ft, fb;
    shading_frame(n, ft, fb);
    ls.direction = normalize3(to_world(V3{sin_theta * cp, sin_theta * sp, 1.0f - one_minus_cos}, ft, fb, n));
    ls.distance = kFloatMax;
    ls.pdf = selection / ((2.0f * kPi) * L.axis_u[3]);
  } else {
    ls.direction = V3{L.axis_u[0], L.axis_u[1], L.axis_u[2]};
    ls.distance = kFloatMax;
    ls.pdf = selection;
    ls.delta = true;
  }
  return ls;
}

// ------------------------------------------------------- environment image
// scene.slang:805-809 (EnvironmentUv): longitude +pi at u = 0 to -pi at 1,
// +Y at v = 0 to -Y at 1.
__device__ __forceinline__ void environment_uv(V3 local, float& u, float& v) {
  const float uu = 0.5f - atan2_pi(local.x, local.z) * 0.5f;
  u = uu - floorf(uu);
  v = acos_pi(fminf(fmaxf(local.y, -1.0f), 1.0f));
}

// sin and cos of pi x for x in [-1, 1].
__device__ __forceinline__ void sincos_pi(float x, float* s, float* c) {
  const float turns = 0.5f * x;
  sincos_turns(turns < 0.0f ? turns + 1.0f : turns, s, c);
}

// scene.slang:811-817 (EnvironmentDirection).
__device__ __forceinline__ V3 environment_direction(float u, float v) {
  float st, ct, sp, cp;
  sincos_pi(v, &st, &ct);
  sincos_pi(1.0f - 2.0f * u, &sp, &cp);
  return V3{st * sp, ct, st * cp};
}

// scene.slang:819-827 (EnvironmentLookup): the bilinear texel, repeating
// across longitudes and clamped at the poles, times the scale.
__device__ __forceinline__ V3 environment_lookup(const SceneView& S, float u, float v) {
  const int W = static_cast<int>(S.env_width), H = static_cast<int>(S.env_height);
  const float x = u * static_cast<float>(W) - 0.5f, y = v * static_cast<float>(H) - 0.5f;
  const float fx0 = floorf(x), fy0 = floorf(y);
  const float fx = x - fx0, fy = y - fy0;
  int i0 = static_cast<int>(fx0) % W;
  if (i0 < 0) i0 += W;
  const int i1 = i0 + 1 == W ? 0 : i0 + 1;
  const int j = static_cast<int>(fy0);
  const int j0 = j < 0 ? 0 : (j > H - 1 ? H - 1 : j), j1 = j + 1 < 0 ? 0 : (j + 1 > H - 1 ? H - 1 : j + 1);
  const float4 t00 = S.env_texels[j0 * W + i0], t10 = S.env_texels[j0 * W + i1];
  const float4 t01 = S.env_texels[j1 * W + i0], t11 = S.env_texels[j1 * W + i1];
  auto lerp2 = [&](float a, float b, float c, float d) {
    return ((1.0f - fx) * a + fx * b) * (1.0f - fy) + ((1.0f - fx) * c + fx * d) * fy;
  };
  const V3 texel{lerp2(t00.x, t10.x, t01.x, t11.x), lerp2(t00.y, t10.y, t01.y, t11.y),
                 lerp2(t00.z, t10.z, t01.z, t11.z)};
  return V3{maxf(texel.x, 0.0f) * S.env_scale[0], maxf(texel.y, 0.0f) * S.env_scale[1],
            maxf(texel.z, 0.0f) * S.env_scale[2]};
}

// scene.slang:829-843 (EnvironmentDirectionPdf): the cell's density per unit
// image area over the 2 pi^2 sin(theta) steradians per unit area.
__device__ __forceinline__ float environment_direction_pdf(const SceneView& S, float u, float v) {
  const uint32_t cu = static_cast<uint32_t>(u * static_cast<float>(S.env_columns));
  const uint32_t cv = static_cast<uint32_t>(v * static_cast<float>(S.env_rows));
  const uint32_t column = cu < S.env_columns - 1 ? cu : S.env_columns - 1;
  const uint32_t row = cv < S.env_rows - 1 ? cv : S.env_rows - 1;
  float st, ct;
  sincos_pi(v, &st, &ct);
  if (!(st > 0.0f)) return 0.0f;
  return S.env_dist[row * S.env_columns + column] / ((2.0f * kPi * kPi) * st);
}

// scene.slang:775-803 (FindInterval, SampleCumulative): a point of [0, 1)
// distributed as the `count` cells of the cumulative distribution at `cdf`.
__device__ __forceinline__ float sample_cumulative(const float* cdf, uint32_t count, float u, uint32_t& cell) {
  uint32_t low = 0, length = count + 1;
  while (length > 0) {
    const uint32_t half = length >> 1, middle = low + half;
    if (cdf[middle] <= u) {
      low = middle + 1;
      length -= half + 1;
    } else {
      length = half;
    }
  }
  cell = (low < 1u ? 1u : (low > count ? count : low)) - 1u;
  const float lo = cdf[cell];
  const float width = cdf[cell + 1] - lo;
  const float offset = width > 0.0f ? (u - lo) / width : 0.5f;
  return fminf((static_cast<float>(cell) + offset) / static_cast<float>(count), 0.99999994f);
}

__device__ __forceinline__ V3 environment_to_world(const SceneView& S, V3 local) {
  const float* f = S.env_frame;
  return normalize3(((local.x * V3{f[0], f[1], f[2]}) + (local.y * V3{f[3], f[4], f[5]})) +
                    (local.z * V3{f[6], f[7], f[8]}));
}

// scene.slang:845-875 (SampleEnvironment): a row by ux, a cell of the row by
// uy, uniformly within the cell.
__device__ __forceinline__ LightSample sample_environment(const SceneView& S, float selection, float ux, float uy) {
  const LightSample none{V3{0.0f, 0.0f, 1.0f}, 0.0f, V3{0.0f, 0.0f, 0.0f}, 0.0f, false};
  const uint32_t cells = S.env_columns * S.env_rows;
  uint32_t row, column;
  const float v = sample_cumulative(S.env_dist + cells, S.env_rows, ux, row);
  const float u = sample_cumulative(S.env_dist + cells + S.env_rows + 1 + row * (S.env_columns + 1), S.env_columns,
                                    uy, column);
  float st, ct;
  sincos_pi(v, &st, &ct);
  const float density = S.env_dist[row * S.env_columns + column];
  if (!(st > 0.0f) || !(density > 0.0f)) return none;
  LightSample ls;
  ls.direction = environment_to_world(S, environment_direction(u, v));
  ls.distance = kFloatMax;
  ls.radiance = environment_lookup(S, u, v);
  ls.pdf = selection * density / ((2.0f * kPi * kPi) * st);
  ls.delta = false;
  return ls;
}

// scene.slang:877-890 (EnvironmentRadiance): the image's radiance along unit
// world `d`, weighed against light sampling when `weigh`.
__device__ __forceinline__ V3 environment_radiance(const SceneView& S, V3 d, bool weigh, float pdf) {
  const float* f = S.env_frame;
  const V3 local{dot3(V3{f[0], f[1], f[2]}, d), dot3(V3{f[3], f[4], f[5]}, d), dot3(V3{f[6], f[7], f[8]}, d)};
  float u, v;
  environment_uv(local, u, v);
  const V3 radiance = environment_lookup(S, u, v);
  if (!weigh || !(S.env_selection > 0.0f)) return radiance;
  return radiance * power_heuristic(pdf, S.env_selection * environment_direction_pdf(S, u, v));
}

// scene.slang:894-909 (SampleLight): the alias table's column and threshold
// choose the light, an emissive triangle or (kAnalytic) an analytic light.
// Kernels of scenes without analytic lights compile the triangle case alone.
template <bool kAnalytic = true, bool kCov = kAnalytic>
__device__ __forceinline__ LightSample sample_light(const SceneView& S, float ux, float uy, float uz, float uw,
                                                    V3 origin) {
  const uint32_t count = S.light_count;
  const uint32_t c = static_cast<uint32_t>(ux * static_cast<float>(count));
  const uint32_t column = c < count - 1 ? c : count - 1;
  const LightRecord rec = S.lights[column];
  const uint32_t light = uy < rec.threshold ? column : rec.alias;
  if constexpr (kAnalytic) {
    const LightRecord chosen = S.lights[light];
    if (chosen.instance == kEnvironmentLight) return sample_environment(S, chosen.pdf, uz, uw);
    if (chosen.instance & kAnalyticLight)
      return sample_analytic(S.analytic[chosen.instance & ~kAnalyticLight], chosen.pdf, uz, uw, origin);
  }
  return sample_triangle_light<kAnalytic, kCov>(S, light, uz, uw, origin);
}

// scene.slang:939-971 (ReconstructReservoirLight): a light-table entry and
// its unit-square coordinates as a sample; a sphere's candidate is a point
// uniformly by area over the whole sphere (its far side contributing 0).
template <bool kAnalytic = true, bool kCov = kAnalytic>
__device__ __forceinline__ LightSample reconstruct_reservoir_light(const SceneView& S, uint32_t index, float u,
                                                                   float v, V3 origin) {
  if constexpr (!kAnalytic) return sample_triangle_light<false, false>(S, index, u, v, origin);
  const LightRecord chosen = S.lights[index];
  if (chosen.instance == kEnvironmentLight) return sample_environment(S, chosen.pdf, u, v);
  if ((chosen.instance & kAnalyticLight) == 0) return sample_triangle_light<true, kCov>(S, index, u, v, origin);
  const AnalyticRecord& L = S.analytic[chosen.instance & ~kAnalyticLight];
  if (L.shape != kSphereLight) return sample_analytic(L, chosen.pdf, u, v, origin);
  const LightSample none{V3{0.0f, 0.0f, 1.0f}, 0.0f, V3{0.0f, 0.0f, 0.0f}, 0.0f, false};
  const float z = 1.0f - 2.0f * u;
  const float rxy = sqrtf(maxf(0.0f, 1.0f - z * z));
  float sp, cp;
  sincos_turns(v, &sp, &cp);
  const V3 normal = V3{rxy * cp, rxy * sp, z};
  const float radius = L.position[3];
  const V3 toward = (V3{L.position[0], L.position[1], L.position[2]} + radius * normal) - origin;
  const float d2 = dot3(toward, toward);
  if (!(d2 > 0.0f)) return none;
  LightSample ls = none;
  ls.distance = sqrtf(d2);
  ls.direction = toward / ls.distance;
  const float cosine = -dot3(normal, ls.direction);
  if (!(cosine > 0.0f)) return none;
  ls.radiance = V3{L.radiance[0], L.radiance[1], L.radiance[2]};
  ls.pdf = chosen.pdf * d2 / ((((cosine * 4.0f) * kPi) * radius) * radius);
  return ls;
}

// scene.slang:973-986 (ReservoirPartitionPdf): the power table's density of
// the direction, a sphere's as its cone would sample it.
template <bool kAnalytic = true>
__device__ __forceinline__ float reservoir_partition_pdf(const SceneView& S, uint32_t index, const LightSample& ls,
                                                         V3 origin) {
  if constexpr (!kAnalytic) return ls.pdf;
  const LightRecord chosen = S.lights[index];
  if ((chosen.instance & kAnalyticLight) && chosen.instance != kEnvironmentLight) {
    const AnalyticRecord& L = S.analytic[chosen.instance & ~kAnalyticLight];
    if (L.shape == kSphereLight) {
      const V3 offset = V3{L.position[0], L.position[1], L.position[2]} - origin;
      return chosen.pdf * sphere_cone_pdf(L.position[3] * L.position[3] / dot3(offset, offset));
    }
  }
  return ls.pdf;
}

// scene.slang:1019-1093 (CrossedLight): the emission of analytic light L if
// the ray crosses it before `distance`, weighed against light sampling by
// the power heuristic when `weigh`.
__device__ __forceinline__ V3 crossed_light(const AnalyticRecord& L, V3 o, V3 d, float distance, bool escaped,
                                            bool weigh, float pdf) {
  const V3 zero{0.0f, 0.0f, 0.0f};
  const uint32_t shape = L.shape;
  float light_pdf = 0.0f;
  if (shape == kSphereLight) {
    // the nearer root without cancellation (Haines et al., Ray Tracing Gems, 2019)
    const float radius = L.position[3];
    const V3 offset = o - V3{L.position[0], L.position[1], L.position[2]};
    const float c = dot3(offset, offset) - radius * radius;
    const float b = -dot3(offset, d);
    const V3 perpendicular = offset + b * d;
    const float disc = radius * radius - dot3(perpendicular, perpendicular);
    if (!(c > 0.0f) || !(b > 0.0f) || disc < 0.0f) return zero;
    const float t = c / (b + sqrtf(disc));
    if (!(t < distance)) return zero;
    light_pdf = sphere_cone_pdf(radius * radius / dot3(offset, offset));
  } else if (shape == kRectLight || shape == kDiskLight) {
    const V3 n = V3{L.normal[0], L.normal[1], L.normal[2]};
    const float facing = dot3(d, n);
    if (!(facing < 0.0f)) return zero;
    const V3 p = V3{L.position[0], L.position[1], L.position[2]};
    const float t = dot3(p - o, n) / facing;
    if (!(t > 0.0f) || !(t < distance)) return zero;
    const V3 q = (o + t * d) - p;
    const V3 u = V3{L.axis_u[0], L.axis_u[1], L.axis_u[2]}, v = V3{L.axis_v[0], L.axis_v[1], L.axis_v[2]};
    const float uu = dot3(u, u), uv = dot3(u, v), vv = dot3(v, v), qu = dot3(q, u), qv = dot3(q, v);
    const float det = uu * vv - uv * uv;
    const float a = (qu * vv - qv * uv) / det;
    const float b = (qv * uu - qu * uv) / det;
    if (shape == kRectLight ? !(fabsf(a) <= 1.0f && fabsf(b) <= 1.0f) : !(a * a + b * b <= 1.0f)) return zero;
    light_pdf = t * t / (-facing * L.normal[3]);
  } else if (shape == kDistantLight && escaped) {
    const V3 toward = V3{L.axis_u[0], L.axis_u[1], L.axis_u[2]};
    const float one_minus_cos = L.axis_u[3];
    const float cosine = dot3(d, toward);
    if (one_minus_cos <= 1.0f) {  // within the cone, by its sine where the cosine is too coarse
      const V3 across = cross3(d, toward);
      if (!(cosine > 0.0f) || !(dot3(across, across) <= one_minus_cos * (2.0f - one_minus_cos))) return zero;
    } else if (!(cosine >= 1.0f - one_minus_cos)) {
      return zero;
    }
    light_pdf = 1.0f / ((2.0f * kPi) * one_minus_cos);
  } else {
    return zero;
  }
  const float weight = weigh ? power_heuristic(pdf, L.radiance[3] * light_pdf) : 1.0f;
  return V3{L.radiance[0], L.radiance[1], L.radiance[2]} * weight;
}

// scene.slang:1095-1108 (OverlapsBox).
__device__ __forceinline__ bool overlaps_box(const float* lo, const float* hi, V3 o, V3 inv, float distance) {
  const float tx0 = (lo[0] - o.x) * inv.x, tx1 = (hi[0] - o.x) * inv.x;
  const float ty0 = (lo[1] - o.y) * inv.y, ty1 = (hi[1] - o.y) * inv.y;
  const float tz0 = (lo[2] - o.z) * inv.z, tz1 = (hi[2] - o.z) * inv.z;
  const float nx = minf(tx0, tx1), ny = minf(ty0, ty1), nz = minf(tz0, tz1);
  const float fx = maxf(tx0, tx1), fy = maxf(ty0, ty1), fz = maxf(tz0, tz1);
  return maxf(maxf(nx, ny), maxf(nz, 0.0f)) <= minf(minf(fx, fy), minf(fz, distance));
}

// scene.slang:1110-1182 (CrossedLights): the emission of every analytic
// light the ray crosses before `distance`: the sphere, rect and disk lights
// through the crossing hierarchy in its order, then, for a ray that
// escaped, the distant lights.
__device__ __forceinline__ V3 crossed_lights(const SceneView& S, V3 o, V3 d, float distance, bool escaped, bool weigh,
                                             float pdf) {
  V3 sum{0.0f, 0.0f, 0.0f};
  if (S.light_node_count != 0) {
    auto safe = [](float x) { return fabsf(x) > 1e-30f ? x : (x < 0.0f ? -1e-30f : 1e-30f); };
    const V3 inv = V3{1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z)};
    uint32_t stack[kLightStack];
    uint32_t depth = 0, node = 0;
    while (true) {
      const LightNode cur = S.light_nodes[node];
      const uint32_t child0 = __float_as_uint(cur.low0[3]), child1 = __float_as_uint(cur.low1[3]);
      const bool enter0 = overlaps_box(cur.low0, cur.high0, o, inv, distance);
      const bool enter1 = child1 != kNoLightChild && overlaps_box(cur.low1, cur.high1, o, inv, distance);
      uint32_t next = 0xFFFFFFFFu;
      if (enter0) {
        if (child0 & kLightLeaf)
          sum = sum + crossed_light(S.analytic[child0 & ~kLightLeaf], o, d, distance, escaped, weigh, pdf);
        else
          next = child0;
      }
      if (enter1) {
        if (child1 & kLightLeaf)
          sum = sum + crossed_light(S.analytic[child1 & ~kLightLeaf], o, d, distance, escaped, weigh, pdf);
        else if (next == 0xFFFFFFFFu)
          next = child1;
        else if (depth < kLightStack)
          stack[depth++] = child1;
      }
      if (next == 0xFFFFFFFFu) {
        if (depth == 0) break;
        next = stack[--depth];
      }
      node = next;
    }
  }
  if (escaped)
    for (uint32_t i = 0; i < S.distant_count; ++i)
      sum = sum + crossed_light(S.analytic[S.distant[i]], o, d, distance, escaped, weigh, pdf);
  return sum;
}

// ------------------------------------------------------------- ReSTIR DI
// A reservoir record (restir.slang:11-17, 48 bytes) as 3 float4: {sample,
// weight_sum, selected_target, M}, {u, v, proposal, flags}, {contribution
// rgb, visibility}. flags: bit 0 the camera ray hit, bit 1 a failed stream.
// Records are stored 4 float4 apart: the reuse record (restir.slang:21-28)
// adds {normalization, supporting candidates, previous index, age}.
constexpr uint32_t kNoReservoirSample = 0xFFFFFFFFu;
constexpr uint32_t kLightRestir = 2u;  // light_sampling: ReSTIR's initial reservoir at the first hit

struct Reservoir {
  uint32_t sample;
  float weight_sum, selected_target;
  uint32_t count;
};

__device__ __forceinline__ Reservoir empty_reservoir() { return Reservoir{kNoReservoirSample, 0.0f, 0.0f, 0u}; }

// reservoir.slang:23-56 (StreamReservoirCandidate).
__device__ __forceinline__ bool stream_candidate(Reservoir& r, uint32_t sample, float target, float proposal,
                                                 float random, bool& selected) {
  selected = false;
  const bool valid_zero = r.weight_sum == 0.0f && r.sample == kNoReservoirSample && r.selected_target == 0.0f;
  const bool valid_positive = r.weight_sum > 0.0f && r.sample != kNoReservoirSample && r.selected_target > 0.0f &&
                              r.count != 0u;
  if (!isfinite(r.weight_sum) || !isfinite(r.selected_target) || !(valid_zero || valid_positive) ||
      !isfinite(target) || target < 0.0f || !isfinite(proposal) || proposal <= 0.0f || !isfinite(random) ||
      random < 0.0f || random >= 1.0f || (target > 0.0f && sample == kNoReservoirSample) || r.count == 0xFFFFFFFFu)
    return false;
  const float weight = target / proposal;
  const float sum = r.weight_sum + weight;
  if (!isfinite(weight) || !isfinite(sum) || (target > 0.0f && weight == 0.0f)) return false;
  selected = weight > 0.0f && (r.weight_sum == 0.0f || random < weight / sum);
  r.weight_sum = sum;
  ++r.count;
  if (selected) {
    r.sample = sample;
    r.selected_target = target;
  }
  return true;
}

// reservoir.slang:58-62.
__device__ __forceinline__ float reservoir_normalization(const Reservoir& r) {
  return r.weight_sum > 0.0f ? (r.weight_sum / static_cast<float>(r.count)) / r.selected_target : 0.0f;
}

// scene.slang:919-943 (ChooseReservoirLight, ReservoirProposalPdf).
__device__ __forceinline__ uint32_t choose_reservoir_light(const SceneView& S, uint32_t mode, float ux, float uy) {
  const uint32_t count = S.light_count;
  const uint32_t c = static_cast<uint32_t>(ux * static_cast<float>(count));
  const uint32_t column = c < count - 1 ? c : count - 1;
  if (mode == 1u) return column;
  if (mode == 2u) {
    if (uy < 0.5f) return column;
    uy = (uy - 0.5f) * 2.0f;
  }
  const LightRecord rec = S.lights[column];
  return uy < rec.threshold ? column : rec.alias;
}
__device__ __forceinline__ float reservoir_proposal_pdf(const SceneView& S, uint32_t mode, uint32_t light) {
  const float power = S.lights[light].pdf;
  const float uniform = 1.0f / static_cast<float>(S.light_count);
  if (mode == 1u) return uniform;
  if (mode == 2u) return 0.5f * (power + uniform);
  return power;
}

// restir.slang:66-74 and 157-161: reflection-side values only; f in the
// light's unit-square measure with the power selection cancelled.
__device__ __forceinline__ V3 candidate_contribution(const Surface& sf, const V3& ft, const V3& fb, V3 wo,
                                                     const LightSample& ls, float selection) {
  if (!(ls.pdf > 0.0f) || dot3(ls.direction, sf.normal) <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
  const V3 n = sf.shading_normal;
  const V3 wi = V3{dot3(ls.direction, ft), dot3(ls.direction, fb), dot3(ls.direction, n)};
  if (!(wi.z > 0.0f)) return V3{0.0f, 0.0f, 0.0f};
  return bsdf_evaluate(sf.bsdf, wo, wi) * ls.radiance * (selection / ls.pdf);
}

// scene.slang:1004-1010.
__device__ __forceinline__ float emitter_pdf(const SceneView& S, uint32_t instance, uint32_t primitive, float area) {
  if (S.light_count == 0) return 0.0f;
  const uint32_t first = S.light_offsets[instance];
  if (first == kNoLights || !(area > 0.0f)) return 0.0f;
  return S.lights[first + primitive].pdf / area;
}

// The outcome of one scattering event.
struct Scatter {
  bool continues;      // path.o/d hold the continuation ray
  bool shadow;         // a shadow ray carries `delta` if it reaches the light
  Ray shadow_ray;
  V3 delta;            // throughput * contribution, added when unoccluded
};

// A shadow ray's contribution in factors, for a ray through covered
// surfaces, whose transmittance multiplies inside the product.
struct ShadowFactors {
  V3 fl;               // f * radiance
  float weight, pdf;   // MIS weight, the light's pdf
  V3 thr;              // the path's throughput at the vertex
};

// ShadowedLight's contribution (transport.slang:109-114) once the shadow
// ray's transmittance is known, times the path's throughput: f * radiance *
// (transmittance * weight / pdf), zero where not finite.
__device__ __forceinline__ V3 shadowed_delta(V3 fl, float weight, float pdf, V3 throughput, float transmittance) {
  const V3 c = fl * (transmittance * weight / pdf);
  return throughput * (finite3(c) ? c : V3{0.0f, 0.0f, 0.0f});
}

// transport.slang:164-170 with EmissionWeight (:150-158): the surface's emission, MIS-weighted against
// light sampling unless the path started here or arrived by a delta or
// transmitted direction, added to the path's radiance.
__device__ __forceinline__ void add_emission(const SceneView& S, const FrameConst& F, PathRegs& p, const Surface& sf,
                                             uint32_t instance, uint32_t primitive) {
  V3 emission = sf.emission;
  if (F.light_sampling != 0 && !(p.last_pdf < 0.0f) && p.depth != 0 && max3(emission) > 0.0f) {
    float w = 1.0f;
    const float area_pdf = emitter_pdf(S, instance, primitive, sf.area);
    if (area_pdf > 0.0f) {
      const V3 toward = sf.position - p.o;
      const float cosine = fabsf(dot3(sf.normal, p.d));
      w = power_heuristic(p.last_pdf, area_pdf * dot3(toward, toward) / cosine);
    }
    emission = emission * w;
  }
  p.radiance = p.radiance + p.throughput * emission;
}

// A receiver surface for reuse (restir.slang:30-62, RestirSurface): what a
// reservoir's target needs at it, kReuseSurfaceWords float4 =
// {origin, view depth}, {geometric normal, instance id low},
// {shading normal, instance id high}, {wo in the shading frame, c0},
// {c1..c4}, {c5..c8}, {c9..c12}, {c13, c14, material id low, high},
// {c15..c18}, {c19, c20, c21, model}. The closure words c: UsdPreviewSurface
// (model 0) keeps {dielectric f90; base colour, alpha; dielectric f0,
// metallic; f0, f90; transmission (negated: thin), eta (negated: smooth)}
// in c0..c14 and zeros after; glTF PBR (model 1) keeps its 22 terms as
// PackGltfPbrTerms orders them (gltf_pbr.slang:159-190) in c0..c21
// (StoreClosure, restir.slang:80-112, with this layout's words).
constexpr uint32_t kReuseSurfaceWords = 10;
__device__ __forceinline__ void store_reuse_surface(float4* out, const Surface& sf, V3 origin, float depth, V3 wo,
                                                    uint64_t instance_id, uint64_t material_id) {
  out[0] = make_float4(origin.x, origin.y, origin.z, depth);
  out[1] = make_float4(sf.normal.x, sf.normal.y, sf.normal.z, __uint_as_float(static_cast<uint32_t>(instance_id)));
  out[2] = make_float4(sf.shading_normal.x, sf.shading_normal.y, sf.shading_normal.z,
                       __uint_as_float(static_cast<uint32_t>(instance_id >> 32)));
  const float mlo = __uint_as_float(static_cast<uint32_t>(material_id));
  const float mhi = __uint_as_float(static_cast<uint32_t>(material_id >> 32));
  if (sf.bsdf.model == kGltfModel) {
    float l[22];
    pack_gltf_terms(sf.bsdf.gltf.t, l);
    out[3] = make_float4(wo.x, wo.y, wo.z, l[0]);
    out[4] = make_float4(l[1], l[2], l[3], l[4]);
    out[5] = make_float4(l[5], l[6], l[7], l[8]);
    out[6] = make_float4(l[9], l[10], l[11], l[12]);
    out[7] = make_float4(l[13], l[14], mlo, mhi);
    out[8] = make_float4(l[15], l[16], l[17], l[18]);
    out[9] = make_float4(l[19], l[20], l[21], 1.0f);
    return;
  }
  const Closure& b = sf.bsdf.preview;
  out[3] = make_float4(wo.x, wo.y, wo.z, b.dielectric_f90);
  out[4] = make_float4(b.base_color.x, b.base_color.y, b.base_color.z, b.alpha);
  out[5] = make_float4(b.dielectric_f0.x, b.dielectric_f0.y, b.dielectric_f0.z, b.metallic);
  out[6] = make_float4(b.f0.x, b.f0.y, b.f0.z, b.f90);
  out[7] = make_float4(b.thin_walled ? -b.transmission : b.transmission, b.smooth ? -b.eta : b.eta, mlo, mhi);
  out[8] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  out[9] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

struct ReuseSurface {
  V3 origin, normal, shading_normal, wo;
  float depth;
  uint64_t instance_id, material_id;
  SurfaceBsdf bsdf;
};

// restir.slang:117-141 (LoadClosure): targets are reflection-side, so the
// preview closure's transmission colour is white; a glTF PBR closure is
// rebuilt from its terms for the stored wo (MakeGltfPbrClosure). Kernels
// of scenes without glTF PBR materials (kGltf false) read the preview
// words only.
template <bool kGltf = false>
__device__ __forceinline__ ReuseSurface load_reuse_surface(const float4* in) {
  const float4 s0 = in[0], s1 = in[1], s2 = in[2], s3 = in[3], s4 = in[4], s5 = in[5], s6 = in[6], s7 = in[7];
  ReuseSurface r;
  r.origin = xyz(s0);
  r.depth = s0.w;
  r.normal = xyz(s1);
  r.shading_normal = xyz(s2);
  r.wo = xyz(s3);
  r.instance_id = static_cast<uint64_t>(__float_as_uint(s1.w)) | (static_cast<uint64_t>(__float_as_uint(s2.w)) << 32);
  r.material_id = static_cast<uint64_t>(__float_as_uint(s7.z)) | (static_cast<uint64_t>(__float_as_uint(s7.w)) << 32);
  if constexpr (kGltf) {
    const float4 s8 = in[8], s9 = in[9];
    if (s9.w != 0.0f) {
      const float l[22] = {s3.w, s4.x, s4.y, s4.z, s4.w, s5.x, s5.y, s5.z, s5.w, s6.x, s6.y,
                           s6.z, s6.w, s7.x, s7.y, s8.x, s8.y, s8.z, s8.w, s9.x, s9.y, s9.z};
      r.bsdf.model = kGltfModel;
      r.bsdf.gltf = make_gltf_closure(unpack_gltf_terms(l), r.wo);
      return r;
    }
  }
  r.bsdf.model = kPreviewModel;
  Closure& b = r.bsdf.preview;
  b.base_color = xyz(s4);
  b.alpha = s4.w;
  b.dielectric_f0 = xyz(s5);
  b.metallic = s5.w;
  b.f0 = xyz(s6);
  b.f90 = s6.w;
  b.dielectric_f90 = s3.w;
  b.transmission = fabsf(s7.x);
  b.thin_walled = s7.x < 0.0f;
  b.eta = fabsf(s7.y);
  b.smooth = s7.y < 0.0f;
  b.transmission_color = V3{1.0f, 1.0f, 1.0f};
  return r;
}

// restir.slang:256-268 (HistoryContribution) for a reconstructed sample.
__device__ __forceinline__ V3 reuse_contribution(const SceneView& S, const ReuseSurface& r, uint32_t light,
                                                 const LightSample& ls) {
  if (!(ls.pdf > 0.0f) || dot3(ls.direction, r.normal) <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
  const V3 n = r.shading_normal;
  const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  const V3 ft = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
  const V3 fb = V3{b, sign + n.y * n.y * a, -n.y};
  const V3 wi = V3{dot3(ls.direction, ft), dot3(ls.direction, fb), dot3(ls.direction, n)};
  if (!(wi.z > 0.0f)) return V3{0.0f, 0.0f, 0.0f};
  return bsdf_evaluate(r.bsdf, r.wo, wi) * ls.radiance * (S.lights[light].pdf / ls.pdf);
}

// restir.slang:235-254 (InitialDirectLight): the first hit's direct light
// from its reservoir record, weighted against BSDF sampling by the power
// heuristic of the light table's density.
template <bool kAnalytic, bool kCov = kAnalytic>
__device__ __forceinline__ V3 initial_direct_light(const SceneView& S, const float4* rec, const Surface& sf,
                                                   const V3& ft, const V3& fb, V3 wo) {
  const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
  const uint32_t light = __float_as_uint(r0.x);
  if ((__float_as_uint(r1.w) & 2u) != 0u || !(r0.y > 0.0f) || light >= S.light_count)
    return V3{0.0f, 0.0f, 0.0f};
  const V3 origin = offset_ray_origin(sf.position, sf.normal);
  const LightSample ls = reconstruct_reservoir_light<kAnalytic, kCov>(S, light, r1.x, r1.y, origin);
  const V3 n = sf.shading_normal;
  const V3 wi = V3{dot3(ls.direction, ft), dot3(ls.direction, fb), dot3(ls.direction, n)};
  const float bsdf = wi.z > 0.0f ? bsdf_pdf(sf.bsdf, wo, wi) : 0.0f;
  const float weight = ls.delta ? 1.0f : power_heuristic(reservoir_partition_pdf<kAnalytic>(S, light, ls, origin), bsdf);
  return xyz(r2) * weight;
}

// transport.slang:235-247 (MissPath) for a continuation ray that leaves the
// scene: the constant environment, then the analytic lights it crossed
// (`crossed`, zero without them), each times the throughput.
template <bool kAnalytic>
__device__ __forceinline__ void escape_terms(const SceneView& S, bool sample_lights, V3 o, V3 d, V3 throughput,
                                             float last_pdf, V3& env, V3& crossed) {
  V3 environment = V3{S.environment[0], S.environment[1], S.environment[2]};
  const bool weigh = sample_lights && !(last_pdf < 0.0f);
  if (kAnalytic && S.env_width != 0) environment = environment + environment_radiance(S, d, weigh, last_pdf);
  env = throughput * environment;
  crossed = V3{0.0f, 0.0f, 0.0f};
  if (kAnalytic && S.analytic_count != 0)
    crossed = throughput * crossed_lights(S, o, d, kFloatMax, true, sample_lights && !(last_pdf < 0.0f), last_pdf);
}

// transport.slang:164-230 (ScatterPath) with DirectLight (:123-135) and
// ShadowedLight (:82-115) up to the shadow ray, which the caller traces.
// `keep`, when given, receives the reconstructed surface; `slot` indexes the
// ReSTIR record of the first hit; `factors`, when given, receives the shadow
// ray's contribution in factors.
// Under BSDF sampling with more than one parallel light, `walk` (kEnabled)
// sums every parallel light's ShadowedLight through shadow rays it walks
// itself (transport.slang:139-146); the default walker has none.
struct NoParallelWalk {
  static constexpr bool kEnabled = false;
  __device__ V3 operator()(const SceneView&, const Surface&, const V3&, const V3&, const V3&) const {
    return V3{0.0f, 0.0f, 0.0f};
  }
};

template <bool kAnalytic, bool kCov = kAnalytic, class Walk = NoParallelWalk,
          bool kGltf = false>  // kCov: covered emitters' coverage; kGltf: glTF PBR surfaces
__device__ __forceinline__ Scatter scatter_path(const SceneView& S, const FrameConst& F, PathRegs& p,
                                                uint32_t tri, uint32_t instance, uint32_t primitive, float u,
                                                float v, Surface* keep = nullptr, uint32_t slot = 0,
                                                ShadowFactors* factors = nullptr, const Walk& walk = Walk{}) {
  Scatter out;
  out.continues = false;
  out.shadow = false;
  const bool sample_lights = F.light_sampling != 0;
  const Surface sf = reconstruct_hit<kAnalytic, kGltf>(S, tri, instance, u, v, p.d);
  if (keep) *keep = sf;
  add_emission(S, F, p, sf, instance, primitive);
  // the analytic lights the continuation ray crossed before the hit
  // (transport.slang:171-178); camera rays do not see them
  if (kAnalytic && p.depth != 0 && S.analytic_count != 0)
    p.radiance = p.radiance + p.throughput * crossed_lights(S, p.o, p.d, dot3(sf.position - p.o, p.d), false,
                                                            sample_lights && !(p.last_pdf < 0.0f), p.last_pdf);
  if (p.depth >= F.max_bounces) return out;
  if (p.depth >= kRouletteBounces) {
    const float survival = minf(max3(p.throughput), kMaxSurvival);
    if (next_random(p.rng) >= survival) return out;
    p.throughput = p.throughput / survival;
  }
  // MakeShadingFrame (sampling.slang:45-56)
  const V3 n = sf.shading_normal;
  const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  const V3 ft = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
  const V3 fb = V3{b, sign + n.y * n.y * a, -n.y};
  const V3 md = -p.d;
  const V3 wo = V3{dot3(md, ft), dot3(md, fb), dot3(md, n)};
  if (F.light_sampling == kLightRestir && p.depth == 0) {
    p.radiance = p.radiance + p.throughput * initial_direct_light<kAnalytic, kCov>(S, F.restir + 4ull * slot, sf, ft, fb, wo);
  } else if (Walk::kEnabled && !sample_lights && S.parallel_count > 1) {
    // ParallelLights (transport.slang:139-146): the lights' sum, then the
    // throughput, as path.radiance += throughput * ParallelLights(...)
    p.radiance = p.radiance + p.throughput * walk(S, sf, ft, fb, wo);
  } else if ((sample_lights && S.light_count != 0) || (kAnalytic && !sample_lights && S.parallel != 0)) {
    // next event estimation (transport.slang:123-135), or under BSDF sampling
    // the one parallel light no sampled direction finds (:139-146), through
    // ShadowedLight (:82-115)
    const V3 origin = offset_ray_origin(sf.position, sf.normal);
    LightSample ls{};
    if (sample_lights) {
      const float ux = next_random(p.rng), uy = next_random(p.rng), uz = next_random(p.rng),
                  uw = next_random(p.rng);
      ls = sample_light<kAnalytic, kCov>(S, ux, uy, uz, uw, origin);
    } else if constexpr (kAnalytic) {
      const AnalyticRecord& L = S.analytic[S.parallel - 1];
      ls = LightSample{V3{L.axis_u[0], L.axis_u[1], L.axis_u[2]}, kFloatMax,
                       V3{L.radiance[0], L.radiance[1], L.radiance[2]}, 1.0f, true};
    }
    if ((ls.pdf > 0.0f) && !(max3(ls.radiance) <= 0.0f) && !(dot3(ls.direction, sf.normal) <= 0.0f)) {
      const V3 wi = V3{dot3(ls.direction, ft), dot3(ls.direction, fb), dot3(ls.direction, n)};
      if (!(wi.z <= 0.0f)) {
        const V3 f = bsdf_evaluate(sf.bsdf, wo, wi);
        if (!(max3(f) <= 0.0f)) {
          const float weight = ls.delta ? 1.0f : power_heuristic(ls.pdf, bsdf_pdf(sf.bsdf, wo, wi));
          out.shadow = true;
          out.shadow_ray = Ray{origin, 0.0f, ls.direction, ls.distance};
          const V3 fl = f * ls.radiance;
          if (factors) *factors = ShadowFactors{fl, weight, ls.pdf, p.throughput};
          out.delta = shadowed_delta(fl, weight, ls.pdf, p.throughput, 1.0f);
        }
      }
    }
  }
  const float ux = next_random(p.rng), uy = next_random(p.rng), uz = next_random(p.rng);
  BsdfSample s;
  if (!bsdf_sample(sf.bsdf, wo, ux, uy, uz, s)) return out;
  const V3 wi = normalize3(((s.wi.x * ft) + (s.wi.y * fb)) + (s.wi.z * n));
  const float side = dot3(wi, sf.normal);
  if (s.transmitted ? side >= 0.0f : side <= 0.0f) return out;
  p.throughput = p.throughput * s.weight;
  p.last_pdf = (s.delta || s.transmitted) ? -1.0f : s.pdf;
  p.o = offset_ray_origin(sf.position, s.transmitted ? -sf.normal : sf.normal);
  p.d = wi;
  ++p.depth;
  out.continues = true;
  return out;
}

}  // namespace lotus
