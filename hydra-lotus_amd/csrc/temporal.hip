// Temporal frames: primary-surface product, reprojection and SVGF, after
// backend/vulkan/shaders/common/temporal.slang (cited per kernel) and its CPU
// oracles backend/vulkan/temporal.cpp. Each kernel is one thread per pixel
// of the full extent (a grid-stride loop); the filter ping-pongs two halves
// of one buffer and reads 16-byte guides (instance id, depth, packed normal)
// instead of the 64-byte surface records on every tap.
#include "temporal.hpp"

namespace lotus {
namespace {

constexpr uint32_t kNoGuide = 0xFFFFFFFFu;  // a guide's depth word for a pixel without history
constexpr uint32_t kVarianceBatches = 4;    // temporal.hpp: kTemporalVarianceBatches

__device__ __forceinline__ uint32_t tid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t tstride() { return gridDim.x * blockDim.x; }

// temporal.slang:95-141 (Luminance, Pack/UnpackAlbedo, Pack/UnpackNormal,
// Demodulate, Remodulate).
__device__ __forceinline__ float luminance(float x, float y, float z) {
  return 0.2126f * x + 0.7152f * y + 0.0722f * z;
}

__device__ __forceinline__ uint32_t pack_albedo(V3 c) {
  const uint32_t r = static_cast<uint32_t>(rintf(minf(maxf(c.x, 0.0f), 1.0f) * 255.0f));
  const uint32_t g = static_cast<uint32_t>(rintf(minf(maxf(c.y, 0.0f), 1.0f) * 255.0f));
  const uint32_t b = static_cast<uint32_t>(rintf(minf(maxf(c.z, 0.0f), 1.0f) * 255.0f));
  return r | (g << 8) | (b << 16);
}

__device__ __forceinline__ V3 unpack_albedo(uint32_t a) {
  return V3{static_cast<float>(a & 0xFFu) / 255.0f, static_cast<float>((a >> 8) & 0xFFu) / 255.0f,
            static_cast<float>((a >> 16) & 0xFFu) / 255.0f};
}

__device__ __forceinline__ uint32_t pack_normal(V3 n) {
  const float s = (fabsf(n.x) + fabsf(n.y)) + fabsf(n.z);
  float ex = n.x / s, ey = n.y / s;
  if (n.z < 0.0f) {
    const float fx = (1.0f - fabsf(ey)) * (ex >= 0.0f ? 1.0f : -1.0f);
    const float fy = (1.0f - fabsf(ex)) * (ey >= 0.0f ? 1.0f : -1.0f);
    ex = fx;
    ey = fy;
  }
  const int qx = static_cast<int>(rintf(minf(maxf(ex, -1.0f), 1.0f) * 32767.0f));
  const int qy = static_cast<int>(rintf(minf(maxf(ey, -1.0f), 1.0f) * 32767.0f));
  return (static_cast<uint32_t>(qx) & 0xFFFFu) | (static_cast<uint32_t>(qy) << 16);
}

__device__ __forceinline__ V3 unpack_normal(uint32_t u) {
  const float x = maxf(static_cast<float>(static_cast<int16_t>(u & 0xFFFFu)) / 32767.0f, -1.0f);
  const float y = maxf(static_cast<float>(static_cast<int16_t>(u >> 16)) / 32767.0f, -1.0f);
  V3 n = V3{x, y, (1.0f - fabsf(x)) - fabsf(y)};
  const float fold = maxf(-n.z, 0.0f);
  n.x += n.x >= 0.0f ? -fold : fold;
  n.y += n.y >= 0.0f ? -fold : fold;
  const float len = sqrtf((n.x * n.x + n.y * n.y) + n.z * n.z);
  return V3{n.x / len, n.y / len, n.z / len};
}

__device__ __forceinline__ V3 demodulate(V3 c, uint32_t albedo) {
  const V3 a = unpack_albedo(albedo);
  return V3{a.x > 0.0f ? c.x / a.x : c.x, a.y > 0.0f ? c.y / a.y : c.y, a.z > 0.0f ? c.z / a.z : c.z};
}

__device__ __forceinline__ V3 remodulate(V3 c, uint32_t albedo) {
  const V3 a = unpack_albedo(albedo);
  return V3{a.x > 0.0f ? c.x * a.x : c.x, a.y > 0.0f ? c.y * a.y : c.y, a.z > 0.0f ? c.z * a.z : c.z};
}

// x^128 by seven squarings (the normal weight).
__device__ __forceinline__ float pow128(float x) {
#pragma unroll
  for (int i = 0; i < 7; ++i) x = x * x;
  return x;
}

__device__ __forceinline__ bool finite4(float4 v) {
  return isfinite(v.x) && isfinite(v.y) && isfinite(v.z) && isfinite(v.w);
}

// temporal.slang:593-627 (WriteTemporalSurface) at the pixel centre.
__global__ void __launch_bounds__(kTraceBlock) k_temporal_surface(SceneView S, FrameConst F, TemporalConst T,
                                                                  TemporalInstances I, TemporalPixel* out,
                                                                  uint2* spill, int spill_cap) {
  __shared__ uint2 ring[kRing * kTraceBlock];
  TraceStack st;
  st.init(ring, spill, spill_cap);
  const uint32_t rw = F.rect[2], total = F.rect[2] * F.rect[3];
  uint32_t visits = 0, tests = 0;
  for (uint32_t i = tid(); i < total; i += tstride()) {
    const uint32_t x = F.rect[0] + i % rw, y = F.rect[1] + i / rw;
    const uint32_t slot = y * F.width + x;
    const float fx = static_cast<float>(x) + 0.5f, fy = static_cast<float>(y) + 0.5f;
    const Ray r = camera_ray(F.clip_to_world, fx * F.ndc_from_pixel[0] + F.ndc_from_pixel[2],
                             fy * F.ndc_from_pixel[1] + F.ndc_from_pixel[3]);
    HitRecord h;
    if (!walk_closest(S.nodes, S.tris, st, r, h, visits, tests, SceneCoverage{&S})) continue;
    const Tri t = S.tris[h.tri];
    const V3 p0 = xyz(t.a), p1 = xyz(t.b), p2 = xyz(t.c);
    const V3 world = (p0 + h.u * (p1 - p0)) + h.v * (p2 - p0);
    float view[4];
    mul_point(T.current_view, world.x, world.y, world.z, 1.0f, view);
    V3 normal = normalize3(cross3(p1 - p0, p2 - p0));
    if (dot3(normal, r.d) > 0.0f) normal = -normal;
    if (!(isfinite(view[0]) && isfinite(view[1]) && isfinite(view[2]) && isfinite(view[3])) || !finite3(normal) ||
        view[3] <= 0.0f || -view[2] / view[3] <= 0.0f)
      continue;
    TemporalPixel o{};
    o.normal_depth = make_float4(normal.x, normal.y, normal.z, -view[2] / view[3]);
    o.instance_id = I.instance_id ? I.instance_id[h.instance] : static_cast<uint64_t>(h.instance) + 1u;
    o.material_id = I.material_id ? I.material_id[h.instance]
                                  : static_cast<uint64_t>(S.instance_material[h.instance]) + 1u;
    const uint32_t iflags = I.flags ? I.flags[h.instance] : 6u;
    o.flags = 1u | iflags;
    const Surface sf = reconstruct_hit(S, h.tri, h.instance, h.u, h.v, r.d);
    // scene.slang:1254-1257, 1396: the model's surface summary
    o.albedo = pack_albedo(surface_albedo(sf));
    o.shading_normal = pack_normal(sf.shading_normal);
    if ((iflags & 2u) != 0u) {
      V3 prev = world;
      if (I.previous_from_current) {
        float pw[4];
        mul_point(I.previous_from_current + 16ull * h.instance, world.x, world.y, world.z, 1.0f, pw);
        prev = V3{pw[0], pw[1], pw[2]};
      }
      float pv[4], clip[4];
      mul_point(T.previous_view, prev.x, prev.y, prev.z, 1.0f, pv);
      mul_point(T.previous_world_to_clip, prev.x, prev.y, prev.z, 1.0f, clip);
      const float px = clip[0] / clip[3], py = clip[1] / clip[3], pz = clip[2] / clip[3];
      const float depth = -pv[2] / pv[3];
      const float qx = T.previous_display[0] + (px * 0.5f + 0.5f) * T.previous_display[2];
      const float qy = T.previous_display[1] + (py * 0.5f + 0.5f) * T.previous_display[3];
      const float bx = static_cast<float>(T.previous_data[0]), by = static_cast<float>(T.previous_data[1]);
      const float ex = bx + static_cast<float>(T.previous_data[2]), ey = by + static_cast<float>(T.previous_data[3]);
      const bool finite = isfinite(pv[0]) && isfinite(pv[1]) && isfinite(pv[2]) && isfinite(pv[3]) &&
                          isfinite(clip[0]) && isfinite(clip[1]) && isfinite(clip[2]) && isfinite(clip[3]) &&
                          isfinite(qx) && isfinite(qy) && isfinite(depth);
      if (finite && pv[3] > 0.0f && clip[3] > 0.0f && depth > 0.0f && pz >= 0.0f && pz <= 1.0f && px >= -1.0f &&
          py >= -1.0f && px < 1.0f && py < 1.0f && qx >= bx && qy >= by && qx < ex && qy < ey) {
        o.motion_depth = make_float4(qx - fx, qy - fy, depth, 1.0f);
        // temporal.slang:611-622 (ReflectionWeight): the share of history to
        // keep as float bits in the reserved word
        if (T.rectification == kRectifyReflection) {
          const V3 to_current = V3{T.current_eye[0], T.current_eye[1], T.current_eye[2]} - world;
          const V3 to_previous = V3{T.previous_eye[0], T.previous_eye[1], T.previous_eye[2]} - prev;
          const float cosine = dot3(to_current, to_previous) / maxf(length3(to_current) * length3(to_previous), 1.0e-20f);
          const float theta = acos_pi(minf(maxf(cosine, -1.0f), 1.0f)) * 3.14159265358979323846f;
          const float lobe = maxf(sf.roughness * sf.roughness, 0.001f);
          const float share = surface_specular_share(sf);
          const float ratio = theta / lobe;
          o.reserved = __float_as_uint((1.0f - share) + share / (1.0f + ratio * ratio));
        }
      }
    }
    out[slot] = o;
  }
}

// temporal.slang:497-554 (ValidateHistory, ReprojectHistory); codes are
// TemporalRejection's (temporal.hpp:169-181).
__device__ __forceinline__ uint32_t validate_history(const TemporalConst& T, const TemporalPixel& c,
                                                     const TemporalPixel& p) {
  if ((p.flags & 1u) == 0u) return 1u;
  const float* th = T.validation;
  const V3 n = xyz(c.normal_depth), pn = xyz(p.normal_depth);
  if (!finite4(c.normal_depth) || !finite4(p.normal_depth) || !finite4(c.motion_depth) ||
      !(isfinite(th[0]) && isfinite(th[1]) && isfinite(th[2]) && isfinite(th[3])) || c.normal_depth.w <= 0.0f ||
      p.normal_depth.w <= 0.0f || c.motion_depth.z <= 0.0f || th[0] < 0.0f || th[1] < 0.0f || th[2] < -1.0f ||
      th[2] > 1.0f || th[3] < 0.0f || (n.x == 0.0f && n.y == 0.0f && n.z == 0.0f) ||
      (pn.x == 0.0f && pn.y == 0.0f && pn.z == 0.0f))
    return 2u;
  if (c.instance_id != p.instance_id) return 3u;
  if (c.material_id != p.material_id) return 4u;
  const float mx = c.motion_depth.x, my = c.motion_depth.y;
  if (sqrtf(mx * mx + my * my) > th[3]) return 5u;
  if (fabsf(p.normal_depth.w - c.motion_depth.z) > th[0] + th[1] * c.motion_depth.z) return 6u;
  if (dot3(normalize3(n), normalize3(pn)) < th[2]) return 7u;
  return 0u;
}

__global__ void __launch_bounds__(256) k_temporal_reproject(TemporalConst T, const TemporalPixel* current,
                                                            const TemporalPixel* previous,
                                                            TemporalValidityPixel* validity) {
  const uint32_t total = T.width * T.height;
  for (uint32_t i = tid(); i < total; i += tstride()) {
    const TemporalPixel c = current[i];
    TemporalValidityPixel o{0u, 0u, 0xFFFFFFFFu, 0u};
    if ((c.flags & 1u) == 0u) o.rejection = 1u;
    else if (T.available == 0u) o.rejection = 8u;
    else if ((c.flags & 2u) == 0u) o.rejection = 9u;
    else if ((c.flags & 4u) == 0u) o.rejection = 4u;
    else if (c.motion_depth.w != 1.0f) o.rejection = 10u;
    else {
      const float ox = (static_cast<float>(i % T.width) + 0.5f) + c.motion_depth.x;
      const float oy = (static_cast<float>(i / T.width) + 0.5f) + c.motion_depth.y;
      if (!(isfinite(ox) && isfinite(oy))) o.rejection = 2u;
      else if (ox < 0.0f || oy < 0.0f || ox >= static_cast<float>(T.width) || oy >= static_cast<float>(T.height))
        o.rejection = 10u;
      else {
        o.previous_index = static_cast<uint32_t>(oy) * T.width + static_cast<uint32_t>(ox);
        o.rejection = validate_history(T, c, previous[o.previous_index]);
      }
    }
    o.valid = o.rejection == 0u ? 1u : 0u;
    validity[i] = o;
  }
}

// temporal.slang:143-150 (RectifyNeighbour): a fresh-batch neighbour inside
// the image, a primary hit on the same placement.
__device__ __forceinline__ bool rectify_neighbour(const TemporalConst& T, const TemporalPixel* surface, int x, int y,
                                                  uint64_t instance, uint32_t& q) {
  if (x < 0 || y < 0 || x >= static_cast<int>(T.width) || y >= static_cast<int>(T.height)) return false;
  q = static_cast<uint32_t>(y) * T.width + static_cast<uint32_t>(x);
  return (surface[q].flags & 1u) != 0u && surface[q].instance_id == instance;
}

// temporal.slang:152-172 (ClampHistory): the reprojected illumination
// clamped per channel to the 3x3 fresh neighbourhood's mean plus or minus one
// standard deviation.
__device__ float4 clamp_history(const TemporalConst& T, const float4* batch, const TemporalPixel* surface, int x,
                                int y, float4 old_color) {
  const TemporalPixel s = surface[static_cast<uint32_t>(y) * T.width + static_cast<uint32_t>(x)];
  V3 sum = V3{0.0f, 0.0f, 0.0f}, squares = V3{0.0f, 0.0f, 0.0f};
  float count = 0.0f;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      uint32_t q;
      if (!rectify_neighbour(T, surface, x + dx, y + dy, s.instance_id, q)) continue;
      const V3 v = demodulate(xyz(batch[q]), surface[q].albedo);
      sum = sum + v;
      squares = squares + v * v;
      count += 1.0f;
    }
  const V3 mean = sum / count;
  const V3 dev = V3{sqrtf(maxf(squares.x / count - mean.x * mean.x, 0.0f)),
                    sqrtf(maxf(squares.y / count - mean.y * mean.y, 0.0f)),
                    sqrtf(maxf(squares.z / count - mean.z * mean.z, 0.0f))};
  const V3 o = demodulate(xyz(old_color), s.albedo);
  const V3 kept = V3{minf(maxf(o.x, mean.x - dev.x), mean.x + dev.x), minf(maxf(o.y, mean.y - dev.y), mean.y + dev.y),
                     minf(maxf(o.z, mean.z - dev.z), mean.z + dev.z)};
  const V3 c = remodulate(kept, s.albedo);
  return make_float4(c.x, c.y, c.z, old_color.w);
}

// temporal.slang:177-212 (AntiLagDrop): the share of history to drop from
// the 5x5 neighbourhood's fresh and reprojected mean luminance against the
// fresh mean's standard error.
__device__ float anti_lag_drop(const TemporalConst& T, const float4* batch, const TemporalPixel* surface,
                               const TemporalValidityPixel* validity, const TemporalRadiancePixel* old_radiance, int x,
                               int y) {
  const uint64_t instance = surface[static_cast<uint32_t>(y) * T.width + static_cast<uint32_t>(x)].instance_id;
  float fresh = 0.0f, squares = 0.0f, count = 0.0f, history = 0.0f, histories = 0.0f;
  for (int dy = -2; dy <= 2; ++dy)
    for (int dx = -2; dx <= 2; ++dx) {
      uint32_t q;
      if (!rectify_neighbour(T, surface, x + dx, y + dy, instance, q)) continue;
      const uint32_t albedo = surface[q].albedo;
      const V3 d = demodulate(xyz(batch[q]), albedo);
      const float l = luminance(d.x, d.y, d.z);
      fresh += l;
      squares += l * l;
      count += 1.0f;
      const TemporalValidityPixel v = validity[q];
      if (v.valid == 0u) continue;
      const TemporalRadiancePixel old = old_radiance[v.previous_index];
      if (old.sample_count == 0u || !finite4(old.color)) continue;
      const V3 h = demodulate(xyz(old.color), albedo);
      history += luminance(h.x, h.y, h.z);
      histories += 1.0f;
    }
  if (histories == 0.0f || count < 2.0f) return 0.0f;
  const float mean = fresh / count;
  const float variance = maxf(squares - count * mean * mean, 0.0f) / (count - 1.0f);
  const float z = fabsf(history / histories - mean) / sqrtf(variance / count + 1.0e-12f);
  return minf(maxf((z - 2.0f) * 0.5f, 0.0f), 1.0f);
}

// temporal.slang:214-217 (ScaleKept): a kept count scaled, to the nearest sample.
__device__ __forceinline__ uint32_t scale_kept(uint32_t kept, float scale) {
  return static_cast<uint32_t>(rintf(static_cast<float>(kept) * minf(maxf(scale, 0.0f), 1.0f)));
}

// temporal.slang:219-267 (AccumulateTemporalRadiance) with the history
// rectification modes.
__global__ void __launch_bounds__(256) k_temporal_accumulate(TemporalConst T, const float4* batch,
                                                             const TemporalPixel* surface,
                                                             const TemporalValidityPixel* validity,
                                                             const TemporalRadiancePixel* old_radiance,
                                                             TemporalRadiancePixel* radiance) {
  const uint32_t total = T.width * T.height;
  for (uint32_t i = tid(); i < total; i += tstride()) {
    TemporalRadiancePixel o{};
    o.color = batch[i];
    if ((surface[i].flags & 1u) != 0u) {
      o.sample_count = T.fresh_samples < T.cap ? T.fresh_samples : T.cap;
      const V3 d = demodulate(xyz(o.color), surface[i].albedo);
      const float l = luminance(d.x, d.y, d.z);
      float m1 = l, m2 = l * l;
      const TemporalValidityPixel v = validity[i];
      if (v.valid != 0u) {
        const TemporalRadiancePixel old = old_radiance[v.previous_index];
        const uint32_t room = T.cap - o.sample_count;
        uint32_t kept = old.sample_count < room ? old.sample_count : room;
        float4 old_color = old.color;
        const bool finite_old = finite4(old.color) && isfinite(old.first_moment) && isfinite(old.second_moment);
        if (kept != 0u && finite_old && T.rectification != 0u) {
          const int x = static_cast<int>(i % T.width), y = static_cast<int>(i / T.width);
          if (T.rectification == kRectifyClamp)
            old_color = clamp_history(T, batch, surface, x, y, old_color);
          else if (T.rectification == kRectifyAntiLag)
            kept = scale_kept(kept, 1.0f - anti_lag_drop(T, batch, surface, validity, old_radiance, x, y));
          else if (T.rectification == kRectifyReflection)
            kept = scale_kept(kept, __uint_as_float(surface[i].reserved));
        }
        if (kept != 0u && finite_old) {
          const uint32_t count = kept + o.sample_count;
          const float blend = static_cast<float>(kept) / static_cast<float>(count);
          o.color.x = o.color.x + (old_color.x - o.color.x) * blend;
          o.color.y = o.color.y + (old_color.y - o.color.y) * blend;
          o.color.z = o.color.z + (old_color.z - o.color.z) * blend;
          o.color.w = o.color.w + (old_color.w - o.color.w) * blend;
          m1 = m1 + (old.first_moment - m1) * blend;
          m2 = m2 + (old.second_moment - m2) * blend;
          o.sample_count = count;
        }
      }
      o.first_moment = m1;
      o.second_moment = m2;
    }
    radiance[i] = o;
  }
}

// temporal.slang:480-493 (SettleTemporalRadiance; CPU temporal.cpp:514-528):
// a settling frame's colour, the still frames' progressive mean `color`
// crossfaded from the last changing frame's colour `prior` over the history
// cap (at most max_samples) where the retained history, read in place since
// nothing moved, has samples.
__global__ void __launch_bounds__(256) k_temporal_settle(TemporalConst T, const float4* prior,
                                                         const TemporalRadiancePixel* old_radiance, float4* color,
                                                         uint32_t samples, uint32_t max_samples) {
  const uint32_t total = T.width * T.height;
  for (uint32_t i = tid(); i < total; i += tstride()) {
    const float4 still = color[i];
    const float4 p = prior[i];
    uint32_t span = old_radiance[i].sample_count != 0u ? T.cap : 0u;
    if (max_samples != 0u && max_samples < span) span = max_samples;
    if (samples >= span || !finite4(p)) continue;
    const float t = static_cast<float>(samples) / static_cast<float>(span);
    color[i] = make_float4(p.x + (still.x - p.x) * t, p.y + (still.y - p.y) * t, p.z + (still.z - p.z) * t,
                           p.w + (still.w - p.w) * t);
  }
}

// A usable neighbour: inside, with a history, on the same placement.
__device__ __forceinline__ bool usable(const TemporalConst& T, const TemporalPixel* surface,
                                       const TemporalRadiancePixel* radiance, int x, int y, uint64_t instance) {
  if (x < 0 || y < 0 || x >= static_cast<int>(T.width) || y >= static_cast<int>(T.height)) return false;
  const uint32_t q = static_cast<uint32_t>(y) * T.width + static_cast<uint32_t>(x);
  return radiance[q].sample_count != 0u && surface[q].instance_id == instance;
}

// The smaller depth step to a usable neighbour along one axis, zero without one.
__device__ __forceinline__ float depth_step(const TemporalConst& T, const TemporalPixel* surface,
                                            const TemporalRadiancePixel* radiance, int x, int y, int dx, int dy,
                                            float depth, uint64_t instance) {
  float step = -1.0f;
  if (usable(T, surface, radiance, x - dx, y - dy, instance))
    step = fabsf(depth - surface[static_cast<uint32_t>(y - dy) * T.width + static_cast<uint32_t>(x - dx)].normal_depth.w);
  if (usable(T, surface, radiance, x + dx, y + dy, instance)) {
    const float other =
        fabsf(depth - surface[static_cast<uint32_t>(y + dy) * T.width + static_cast<uint32_t>(x + dx)].normal_depth.w);
    step = step < 0.0f ? other : minf(step, other);
  }
  return maxf(step, 0.0f);
}

// temporal.slang:313-370 (EstimateTemporalVariance; CPU temporal.cpp:375-421).
// A workgroup takes 16 x 16 pixels and first stages what the 7 x 7 spatial
// estimate reads of its pixels and a 3-pixel apron (usable flag, instance id,
// normal and depth, moments) in LDS, so the 49 taps of each pixel read LDS
// rather than 96 bytes of records each through the L1. The taps run in the
// same order with the same arithmetic.
constexpr int kVarTile = 16, kVarApron = 3, kVarSpan = kVarTile + 2 * kVarApron;

__global__ void __launch_bounds__(256) k_temporal_variance(TemporalConst T, const TemporalPixel* surface,
                                                           TemporalRadiancePixel* radiance, float4* filter) {
  __shared__ float4 nd[kVarSpan * kVarSpan];
  __shared__ float2 mom[kVarSpan * kVarSpan];
  __shared__ unsigned long long inst[kVarSpan * kVarSpan];
  __shared__ uint32_t ok[kVarSpan * kVarSpan];
  const uint32_t total = T.width * T.height;
  const int W = static_cast<int>(T.width), H = static_cast<int>(T.height);
  const float batch = static_cast<float>(T.fresh_samples < T.cap ? T.fresh_samples : T.cap);
  const uint32_t tiles_x = (T.width + kVarTile - 1) / kVarTile;
  const uint32_t tiles = tiles_x * ((T.height + kVarTile - 1) / kVarTile);
  const int lx = static_cast<int>(threadIdx.x % kVarTile), ly = static_cast<int>(threadIdx.x / kVarTile);
  for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const int x0 = static_cast<int>(t % tiles_x) * kVarTile - kVarApron;
    const int y0 = static_cast<int>(t / tiles_x) * kVarTile - kVarApron;
    __syncthreads();  // the previous tile's taps are done
    for (int k = static_cast<int>(threadIdx.x); k < kVarSpan * kVarSpan; k += static_cast<int>(blockDim.x)) {
      const int gx = x0 + k % kVarSpan, gy = y0 + k / kVarSpan;
      uint32_t usable_here = 0u;
      if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
        const uint32_t q = static_cast<uint32_t>(gy) * T.width + static_cast<uint32_t>(gx);
        usable_here = radiance[q].sample_count != 0u ? 1u : 0u;
        mom[k] = make_float2(radiance[q].first_moment, radiance[q].second_moment);
        nd[k] = surface[q].normal_depth;
        inst[k] = surface[q].instance_id;
      }
      ok[k] = usable_here;
    }
    __syncthreads();
    const int x = x0 + kVarApron + lx, y = y0 + kVarApron + ly;
    if (x >= W || y >= H) continue;
    const uint32_t i = static_cast<uint32_t>(y) * T.width + static_cast<uint32_t>(x);
    const int centre = (ly + kVarApron) * kVarSpan + (lx + kVarApron);
    // usable(): a pixel of the image with history on the same instance
    const auto usable = [&](int k, uint64_t instance) { return ok[k] != 0u && inst[k] == instance; };
    const TemporalRadiancePixel c = radiance[i];
    float variance = 0.0f;
    const TemporalPixel s = surface[i];
    if (c.sample_count != 0u) {
      const float batches = static_cast<float>(c.sample_count) / batch;
      if (batches >= static_cast<float>(kVarianceBatches)) {
        variance = maxf(0.0f, c.second_moment - c.first_moment * c.first_moment);
      } else {
        const float depth = s.normal_depth.w;
        // depth_step(): the smaller depth step to a usable neighbour along one axis
        const auto step = [&](int off) {
          float st = -1.0f;
          if (usable(centre - off, s.instance_id)) st = fabsf(depth - nd[centre - off].w);
          if (usable(centre + off, s.instance_id)) {
            const float other = fabsf(depth - nd[centre + off].w);
            st = st < 0.0f ? other : minf(st, other);
          }
          return maxf(st, 0.0f);
        };
        const float phi = maxf(step(1) + step(kVarSpan), 1.0e-8f) * 3.0f;
        float ws = 0.0f, s1 = 0.0f, s2 = 0.0f;
        for (int dy = -3; dy <= 3; ++dy) {
          for (int dx = -3; dx <= 3; ++dx) {
            const int k = centre + dy * kVarSpan + dx;
            if (!usable(k, s.instance_id)) continue;
            const float4 n = nd[k];
            const float2 nb = mom[k];
            const float distance = sqrtf(static_cast<float>(dx * dx + dy * dy));
            const float wz = dx == 0 && dy == 0 ? 0.0f : fabsf(depth - n.w) / (phi * distance);
            const float cosine = (s.normal_depth.x * n.x + s.normal_depth.y * n.y) + s.normal_depth.z * n.z;
            const float wn = pow128(minf(maxf(cosine, 0.0f), 1.0f));
            const float wl = fabsf(c.first_moment - nb.x) / 10.0f;
            const float w = expf(-(wl + wz)) * wn;
            s1 += w * nb.x;
            s2 += w * nb.y;
            ws += w;
          }
        }
        ws = maxf(ws, 1.0e-6f);
        const float m1 = s1 / ws, m2 = s2 / ws;
        variance = maxf(0.0f, m2 - m1 * m1) * (static_cast<float>(kVarianceBatches) / batches);
      }
    }
    radiance[i].variance = variance;
    if (T.iterations == 0u) continue;
    // iteration 0's input: the illumination and the variance of its mean
    float4 in = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint4 guide = make_uint4(0u, 0u, kNoGuide, 0u);
    if (c.sample_count != 0u) {
      const float batches = static_cast<float>(c.sample_count) / batch;
      const V3 d = demodulate(xyz(c.color), s.albedo);
      in = make_float4(d.x, d.y, d.z, variance / batches);
      guide = make_uint4(static_cast<uint32_t>(s.instance_id), static_cast<uint32_t>(s.instance_id >> 32),
                         __float_as_uint(s.normal_depth.w), s.shading_normal);
    }
    filter[i] = in;
    reinterpret_cast<uint4*>(filter)[2ull * total + i] = guide;
    // the guide's shading normal unpacked once, for every tap of every iteration
    const V3 un = c.sample_count != 0u ? unpack_normal(s.shading_normal) : V3{0.0f, 0.0f, 0.0f};
    filter[3ull * total + i] = make_float4(un.x, un.y, un.z, 0.0f);
  }
}

__device__ __forceinline__ bool filter_neighbour(const TemporalConst& T, const uint4* guides, int x, int y, uint2 inst,
                                                 uint4& g) {
  g = make_uint4(0u, 0u, kNoGuide, 0u);
  if (x < 0 || y < 0 || x >= static_cast<int>(T.width) || y >= static_cast<int>(T.height)) return false;
  g = guides[static_cast<uint32_t>(y) * T.width + static_cast<uint32_t>(x)];
  return g.z != kNoGuide && g.x == inst.x && g.y == inst.y;
}

__device__ __forceinline__ float filter_depth_step(const TemporalConst& T, const uint4* guides, int x, int y, int dx,
                                                   int dy, float depth, uint2 inst) {
  float step = -1.0f;
  uint4 g;
  if (filter_neighbour(T, guides, x - dx, y - dy, inst, g)) step = fabsf(depth - __uint_as_float(g.z));
  if (filter_neighbour(T, guides, x + dx, y + dy, inst, g)) {
    const float other = fabsf(depth - __uint_as_float(g.z));
    step = step < 0.0f ? other : minf(step, other);
  }
  return maxf(step, 0.0f);
}

// temporal.slang:405-454 (FilterIteration; CPU temporal.cpp:424-512): one
// a-trous iteration at pixel (x, y) over input half `in`.
__device__ float4 filter_iteration(const TemporalConst& T, const float4* in, const uint4* guides,
                                   const float4* normals, int x, int y, uint32_t iteration) {
  const uint32_t index = static_cast<uint32_t>(y) * T.width + static_cast<uint32_t>(x);
  const uint4 center = guides[index];
  if (center.z == kNoGuide) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint2 inst = make_uint2(center.x, center.y);
  uint4 g;
  float gaussian = 0.0f, gaussian_weight = 0.0f;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      if (!filter_neighbour(T, guides, x + dx, y + dy, inst, g)) continue;
      const float k = (dx == 0 ? 0.5f : 0.25f) * (dy == 0 ? 0.5f : 0.25f);
      gaussian += k * in[static_cast<uint32_t>(y + dy) * T.width + static_cast<uint32_t>(x + dx)].w;
      gaussian_weight += k;
    }
  const float phi_l = 4.0f * sqrtf(maxf(gaussian / gaussian_weight, 0.0f) + 1.0e-10f);
  const float depth = __uint_as_float(center.z);
  const float phi_z = maxf(filter_depth_step(T, guides, x, y, 1, 0, depth, inst) +
                               filter_depth_step(T, guides, x, y, 0, 1, depth, inst),
                           1.0e-8f);
  const V3 normal = xyz(normals[index]);
  const float4 cv = in[index];
  const float lum = luminance(cv.x, cv.y, cv.z);
  const int step = 1 << iteration;
  const float kernel[3] = {0.375f, 0.25f, 0.0625f};
  float sx = 0.0f, sy = 0.0f, sz = 0.0f, variance = 0.0f, ws = 0.0f;
#pragma unroll
  for (int ty = -2; ty <= 2; ++ty)
#pragma unroll
    for (int tx = -2; tx <= 2; ++tx) {
      const int qx = x + tx * step, qy = y + ty * step;
      if (!filter_neighbour(T, guides, qx, qy, inst, g)) continue;
      const float4 v = in[static_cast<uint32_t>(qy) * T.width + static_cast<uint32_t>(qx)];
      float w = kernel[tx < 0 ? -tx : tx] * kernel[ty < 0 ? -ty : ty];
      if (tx != 0 || ty != 0) {
        const float distance = static_cast<float>(step) * sqrtf(static_cast<float>(tx * tx + ty * ty));
        const float wz = fabsf(depth - __uint_as_float(g.z)) / (phi_z * distance);
        const float wl = fabsf(lum - luminance(v.x, v.y, v.z)) / phi_l;
        const V3 other = xyz(normals[static_cast<uint32_t>(qy) * T.width + static_cast<uint32_t>(qx)]);
        const float cosine = (normal.x * other.x + normal.y * other.y) + normal.z * other.z;
        const float wn = pow128(minf(maxf(cosine, 0.0f), 1.0f));
        w = w * expf(-(wl + wz)) * wn;
      }
      sx += w * v.x;
      sy += w * v.y;
      sz += w * v.z;
      variance += w * w * v.w;
      ws += w;
    }
  return make_float4(sx / ws, sy / ws, sz / ws, variance / (ws * ws));
}

// One a-trous iteration over LDS tiles, for steps of at most 4: a 16 x 16
// block of pixels stages the guides and inputs of the block and its apron
// of 2 s pixels (out of the image: no guide), (16 + 4 s)^2 entries of 32
// bytes, so the 3 x 3 variance Gaussian, the depth steps and the 25 taps read
// LDS instead of 16-byte global loads. The arithmetic is filter_iteration's,
// in the same order, so every result is the same. (Larger steps would need
// more LDS than a CU holds; they use k_temporal_filter.) kComposite: the
// last iteration, written into the colour as k_temporal_composite writes it.
template <int kStep, bool kComposite>
__global__ void __launch_bounds__(256) k_temporal_filter_tiled(TemporalConst T, float4* filter,
                                                               const TemporalPixel* surface,
                                                               const TemporalRadiancePixel* radiance, float4* color) {
  constexpr int kTile = 16, kApron = 2 * kStep, kSpan = kTile + 2 * kApron;
  constexpr uint32_t kIteration = kStep == 1 ? 0u : kStep == 2 ? 1u : 2u;
  __shared__ uint4 lg[kSpan * kSpan];
  __shared__ float4 lv[kSpan * kSpan];
  __shared__ float4 ln[kSpan * kSpan];  // each entry's unpacked normal and luminance
  const uint32_t total = T.width * T.height;
  const int W = static_cast<int>(T.width), H = static_cast<int>(T.height);
  const float4* in = filter + static_cast<size_t>(kIteration & 1u) * total;
  float4* out = filter + static_cast<size_t>((kIteration + 1u) & 1u) * total;
  const uint4* guides = reinterpret_cast<const uint4*>(filter) + 2ull * total;
  const float4* normals = filter + 3ull * total;
  const uint32_t tiles_x = (T.width + kTile - 1) / kTile;
  const uint32_t tiles = tiles_x * ((T.height + kTile - 1) / kTile);
  const int lx = static_cast<int>(threadIdx.x % kTile), ly = static_cast<int>(threadIdx.x / kTile);
  for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const int x0 = static_cast<int>(t % tiles_x) * kTile - kApron;
    const int y0 = static_cast<int>(t / tiles_x) * kTile - kApron;
    __syncthreads();  // the previous tile's taps are done
    for (int k = static_cast<int>(threadIdx.x); k < kSpan * kSpan; k += static_cast<int>(blockDim.x)) {
      const int gx = x0 + k % kSpan, gy = y0 + k / kSpan;
      if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
        const uint32_t q = static_cast<uint32_t>(gy) * T.width + static_cast<uint32_t>(gx);
        const float4 v = in[q];
        const float4 n = normals[q];
        lg[k] = guides[q];
        lv[k] = v;
        ln[k] = make_float4(n.x, n.y, n.z, luminance(v.x, v.y, v.z));
      } else {
        lg[k] = make_uint4(0u, 0u, kNoGuide, 0u);
        lv[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ln[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
    }
    __syncthreads();
    const int x = x0 + kApron + lx, y = y0 + kApron + ly;
    if (x >= W || y >= H) continue;
    const uint32_t index = static_cast<uint32_t>(y) * T.width + static_cast<uint32_t>(x);
    if constexpr (kComposite) {
      const TemporalRadiancePixel r = radiance[index];
      if (r.sample_count == 0u || T.iterations == 0u) {
        color[index] = r.color;
        continue;
      }
    }
    const int centre = (ly + kApron) * kSpan + (lx + kApron);
    const uint4 center = lg[centre];
    float4 result = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (center.z != kNoGuide) {
      const uint2 inst = make_uint2(center.x, center.y);
      const auto match = [&](const uint4& g) { return g.z != kNoGuide && g.x == inst.x && g.y == inst.y; };
      float gaussian = 0.0f, gaussian_weight = 0.0f;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int k = centre + dy * kSpan + dx;
          if (!match(lg[k])) continue;
          const float kw = (dx == 0 ? 0.5f : 0.25f) * (dy == 0 ? 0.5f : 0.25f);
          gaussian += kw * lv[k].w;
          gaussian_weight += kw;
        }
      const float phi_l = 4.0f * sqrtf(maxf(gaussian / gaussian_weight, 0.0f) + 1.0e-10f);
      const float depth = __uint_as_float(center.z);
      // filter_depth_step along x, then y
      const auto depth_step = [&](int off) {
        float st = -1.0f;
        const uint4 a = lg[centre - off], b = lg[centre + off];
        if (match(a)) st = fabsf(depth - __uint_as_float(a.z));
        if (match(b)) {
          const float other = fabsf(depth - __uint_as_float(b.z));
          st = st < 0.0f ? other : minf(st, other);
        }
        return maxf(st, 0.0f);
      };
      const float phi_z = maxf(depth_step(1) + depth_step(kSpan), 1.0e-8f);
      const float4 cn = ln[centre];
      const V3 normal = V3{cn.x, cn.y, cn.z};
      const float lum = cn.w;
      const float kernel[3] = {0.375f, 0.25f, 0.0625f};
      float sx = 0.0f, sy = 0.0f, sz = 0.0f, variance = 0.0f, ws = 0.0f;
#pragma unroll
      for (int ty = -2; ty <= 2; ++ty)
#pragma unroll
        for (int tx = -2; tx <= 2; ++tx) {
          const int k = centre + ty * kStep * kSpan + tx * kStep;
          const uint4 g = lg[k];
          if (!match(g)) continue;
          const float4 v = lv[k];
          float w = kernel[tx < 0 ? -tx : tx] * kernel[ty < 0 ? -ty : ty];
          if (tx != 0 || ty != 0) {
            const float4 on = ln[k];
            const float distance = static_cast<float>(kStep) * sqrtf(static_cast<float>(tx * tx + ty * ty));
            const float wz = fabsf(depth - __uint_as_float(g.z)) / (phi_z * distance);
            const float wl = fabsf(lum - on.w) / phi_l;
            const V3 other = V3{on.x, on.y, on.z};
            const float cosine = (normal.x * other.x + normal.y * other.y) + normal.z * other.z;
            const float wn = pow128(minf(maxf(cosine, 0.0f), 1.0f));
            w = w * expf(-(wl + wz)) * wn;
          }
          sx += w * v.x;
          sy += w * v.y;
          sz += w * v.z;
          variance += w * w * v.w;
          ws += w;
        }
      result = make_float4(sx / ws, sy / ws, sz / ws, variance / (ws * ws));
    }
    if constexpr (kComposite) {
      const V3 c = remodulate(V3{result.x, result.y, result.z}, surface[index].albedo);
      color[index] = make_float4(c.x, c.y, c.z, radiance[index].color.w);
    } else {
      out[index] = result;
    }
  }
}

__global__ void __launch_bounds__(256) k_temporal_filter(TemporalConst T, float4* filter, uint32_t iteration) {
  const uint32_t total = T.width * T.height;
  const float4* in = filter + static_cast<size_t>(iteration & 1u) * total;
  float4* out = filter + static_cast<size_t>((iteration + 1u) & 1u) * total;
  const uint4* guides = reinterpret_cast<const uint4*>(filter) + 2ull * total;
  const float4* normals = filter + 3ull * total;
  for (uint32_t i = tid(); i < total; i += tstride())
    out[i] = filter_iteration(T, in, guides, normals, static_cast<int>(i % T.width), static_cast<int>(i / T.width),
                              iteration);
}

// temporal.slang:468-478 (CompositeTemporalRadiance): the last iteration,
// remodulated, with the accumulated alpha; pixels without a history (and
// every pixel without filtering) keep the accumulated colour.
__global__ void __launch_bounds__(256) k_temporal_composite(TemporalConst T, const TemporalPixel* surface,
                                                            const TemporalRadiancePixel* radiance,
                                                            const float4* filter, float4* color) {
  const uint32_t total = T.width * T.height;
  const uint32_t last = T.iterations ? T.iterations - 1u : 0u;
  const float4* in = filter + static_cast<size_t>(last & 1u) * total;
  const uint4* guides = reinterpret_cast<const uint4*>(filter) + 2ull * total;
  const float4* normals = filter + 3ull * total;
  for (uint32_t i = tid(); i < total; i += tstride()) {
    const TemporalRadiancePixel r = radiance[i];
    if (r.sample_count == 0u || T.iterations == 0u) {
      color[i] = r.color;
      continue;
    }
    const float4 f =
        filter_iteration(T, in, guides, normals, static_cast<int>(i % T.width), static_cast<int>(i / T.width), last);
    const V3 c = remodulate(V3{f.x, f.y, f.z}, surface[i].albedo);
    color[i] = make_float4(c.x, c.y, c.z, r.color.w);
  }
}

uint32_t pixel_blocks(uint32_t total) {
  const uint32_t b = (total + 255u) / 256u;
  return b < 4096u ? (b ? b : 1u) : 4096u;
}

}  // namespace

void launch_temporal_surface(const Launch& L, const SceneView& S, const FrameConst& F, const TemporalConst& T,
                             const TemporalInstances& I, TemporalPixel* out) {
  k_temporal_surface<<<L.grid.camera_pass, kTraceBlock, 0, L.stream>>>(S, F, T, I, out, L.spill, L.spill_cap);
}

void launch_temporal_reproject(const Launch& L, const TemporalConst& T, const TemporalPixel* current,
                               const TemporalPixel* previous, TemporalValidityPixel* validity) {
  k_temporal_reproject<<<pixel_blocks(T.width * T.height), 256, 0, L.stream>>>(T, current, previous, validity);
}

void launch_temporal_accumulate(const Launch& L, const TemporalConst& T, const float4* batch,
                                const TemporalPixel* surface, const TemporalValidityPixel* validity,
                                const TemporalRadiancePixel* old_radiance, TemporalRadiancePixel* radiance) {
  k_temporal_accumulate<<<pixel_blocks(T.width * T.height), 256, 0, L.stream>>>(T, batch, surface, validity,
                                                                               old_radiance, radiance);
}

void launch_temporal_settle(const Launch& L, const TemporalConst& T, const float4* prior,
                            const TemporalRadiancePixel* old_radiance, float4* color, uint32_t samples,
                            uint32_t max_samples) {
  k_temporal_settle<<<pixel_blocks(T.width * T.height), 256, 0, L.stream>>>(T, prior, old_radiance, color, samples,
                                                                           max_samples);
}

void launch_temporal_variance(const Launch& L, const TemporalConst& T, const TemporalPixel* surface,
                              TemporalRadiancePixel* radiance, float4* filter) {
  k_temporal_variance<<<pixel_blocks(T.width * T.height), 256, 0, L.stream>>>(T, surface, radiance, filter);
}

void launch_temporal_filter(const Launch& L, const TemporalConst& T, const TemporalPixel* surface,
                            const TemporalRadiancePixel* radiance, float4* filter, float4* color) {
  const uint32_t blocks = pixel_blocks(T.width * T.height);
  // iterations 0-2 (steps 1, 2, 4) over LDS tiles, later ones from global memory
  const auto iteration = [&](uint32_t it, bool composite) {
    if (it == 0 && composite)
      k_temporal_filter_tiled<1, true><<<blocks, 256, 0, L.stream>>>(T, filter, surface, radiance, color);
    else if (it == 0)
      k_temporal_filter_tiled<1, false><<<blocks, 256, 0, L.stream>>>(T, filter, surface, radiance, color);
    else if (it == 1 && composite)
      k_temporal_filter_tiled<2, true><<<blocks, 256, 0, L.stream>>>(T, filter, surface, radiance, color);
    else if (it == 1)
      k_temporal_filter_tiled<2, false><<<blocks, 256, 0, L.stream>>>(T, filter, surface, radiance, color);
    else if (it == 2 && composite)
      k_temporal_filter_tiled<4, true><<<blocks, 256, 0, L.stream>>>(T, filter, surface, radiance, color);
    else if (it == 2)
      k_temporal_filter_tiled<4, false><<<blocks, 256, 0, L.stream>>>(T, filter, surface, radiance, color);
    else if (composite)
      k_temporal_composite<<<blocks, 256, 0, L.stream>>>(T, surface, radiance, filter, color);
    else
      k_temporal_filter<<<blocks, 256, 0, L.stream>>>(T, filter, it);
  };
  if (T.iterations == 0u) {
    k_temporal_composite<<<blocks, 256, 0, L.stream>>>(T, surface, radiance, filter, color);
    return;
  }
  for (uint32_t it = 0; it + 1u < T.iterations; ++it) iteration(it, false);
  iteration(T.iterations - 1u, true);
}

}  // namespace lotus
