// The wavefront integrator's kernels for gfx950 (the reference's Camera /
// Intersect / Shade / Tail kernels, backend/vulkan/shaders/wavefront/
// wavefront.slang:206-309, and its resolve pass, path_trace.slang:106-139),
// plus the pixel-centre camera pass (path_trace.slang:69-97) and a plain
// trace kernel over caller rays.
//
// Scheduling differs from the reference by design: HIP has no indirect
// dispatch, so every kernel is a persistent grid (a few workgroups per CU)
// striding over a queue whose length it reads from device memory. Queues are
// structure-of-arrays rays and hits in queue order, appended with one atomic
// per wave (ballot + popcount prefix), so a kernel's loads of its queue are
// contiguous. Next event estimation's shadow rays get their own queue and
// kernel, between shading and the next intersection, so the radiance sums
// keep the reference's order.
#include "kernels.hpp"

namespace lotus {
namespace {

template <int N>
__device__ __forceinline__ TraversalStack<N> make_stack(uint2* spill, int cap) {
  TraversalStack<N> st;
  st.init(spill, gridDim.x * blockDim.x, blockIdx.x * blockDim.x + threadIdx.x, cap);
  return st;
}

// Adds a lane's traversal counts to stats[2 kind], stats[2 kind + 1]
// (wave sum, one atomic per wave).
template <int N>
__device__ __forceinline__ void flush_stats(unsigned long long* stats, int kind, const TraversalStack<N>& st) {
  uint32_t v = st.visits, t = st.tests;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v += __shfl_xor(v, off);
    t += __shfl_xor(t, off);
  }
  if (__lane_id() == 0) {
    atomicAdd(stats + 2 * kind, static_cast<unsigned long long>(v));
    atomicAdd(stats + 2 * kind + 1, static_cast<unsigned long long>(t));
  }
}

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t gstride() { return gridDim.x * blockDim.x; }

// Counters per bounce b: [4b] ray-queue entries (camera rays at b = 0),
// [4b+1] hit-queue entries, [4b+2] shadow-queue entries.
__device__ __forceinline__ uint32_t* ray_count(uint32_t* c, uint32_t b) { return c + 4 * b; }
__device__ __forceinline__ uint32_t* hit_count(uint32_t* c, uint32_t b) { return c + 4 * b + 1; }
__device__ __forceinline__ uint32_t* shadow_count(uint32_t* c, uint32_t b) { return c + 4 * b + 2; }

__device__ __forceinline__ void push_hit(const Queues& q, uint32_t k, uint32_t slot, V3 o, V3 d,
                                         const HitRecord& h) {
  q.hit_o[k] = make_float4(o.x, o.y, o.z, __uint_as_float(slot));
  q.hit_d[k] = make_float4(d.x, d.y, d.z, __uint_as_float(h.tri));
  q.hit_h[k] = make_uint4(h.instance, h.primitive, __float_as_uint(h.u), __float_as_uint(h.v));
}

// wavefront.slang:206-225: a camera path per pixel of the data window, its
// camera ray traced at once; a hit is queued for shading, a miss leaves no
// sample.
__global__ void __launch_bounds__(kTraceBlock) k_camera(SceneView S, FrameConst F, Queues q, uint32_t* counters,
                                                        uint2* spill, int spill_cap, unsigned long long* stats) {
  const uint32_t rw = F.rect[2], total = F.rect[2] * F.rect[3];
  const Traverser<kStackRegs> tr{S.nodes, S.tris};
  TraversalStack<kStackRegs> st = make_stack<kStackRegs>(spill, spill_cap);
  for (uint32_t i = gtid(); i < total; i += gstride()) {
    const uint32_t x = F.rect[0] + i % rw, y = F.rect[1] + i / rw;
    const uint32_t slot = y * F.width + x;
    uint32_t rng;
    const Ray r = start_path(F, x, y, F.sample_index, rng);
    HitRecord h;
    const bool hit = tr.closest(r, st, h);
    wave_count(ray_count(counters, 0), true);
    const uint32_t k = wave_append(hit_count(counters, 0), hit);
    if (hit) {
      push_hit(q, k, slot, r.o, r.d, h);
      q.rad[slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(rng));
    } else {
      q.rad[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
  flush_stats(stats, 0, st);
}

// wavefront.slang:230-262: continuation rays of bounce b; a miss ends the
// path with the environment (transport.slang:436-448).
__global__ void __launch_bounds__(kTraceBlock) k_intersect(SceneView S, FrameConst F, Queues q, uint32_t* counters,
                                                           uint2* spill, int spill_cap, uint32_t bounce, unsigned long long* stats) {
  const uint32_t count = *ray_count(counters, bounce);
  const Traverser<kStackRegs> tr{S.nodes, S.tris};
  TraversalStack<kStackRegs> st = make_stack<kStackRegs>(spill, spill_cap);
  const V3 env = V3{S.environment[0], S.environment[1], S.environment[2]};
  for (uint32_t i = gtid(); i < count; i += gstride()) {
    const float4 o4 = q.ray_o[i], d4 = q.ray_d[i];
    const uint32_t slot = __float_as_uint(o4.w);
    const Ray r{xyz(o4), 0.0f, xyz(d4), d4.w};
    HitRecord h;
    const bool hit = tr.closest(r, st, h);
    const uint32_t k = wave_append(hit_count(counters, bounce), hit);
    if (hit) {
      push_hit(q, k, slot, r.o, r.d, h);
    } else {
      const float4 t4 = q.thr[slot], r4 = q.rad[slot];
      const V3 rad = xyz(r4) + xyz(t4) * env;
      q.rad[slot] = make_float4(rad.x, rad.y, rad.z, 1.0f);
    }
  }
  flush_stats(stats, 1, st);
}

// wavefront.slang:264-282: one scattering event per queued hit.
__global__ void __launch_bounds__(kShadeBlock) k_shade(SceneView S, FrameConst F, Queues q, uint32_t* counters,
                                                       uint32_t bounce) {
  const uint32_t count = *hit_count(counters, bounce);
  for (uint32_t i = gtid(); i < count; i += gstride()) {
    const float4 o4 = q.hit_o[i], d4 = q.hit_d[i];
    const uint4 h4 = q.hit_h[i];
    const uint32_t slot = __float_as_uint(o4.w);
    PathRegs p;
    p.o = xyz(o4);
    p.d = xyz(d4);
    p.depth = bounce;
    if (bounce == 0) {
      p.throughput = V3{1.0f, 1.0f, 1.0f};
      p.last_pdf = 0.0f;
    } else {
      const float4 t4 = q.thr[slot];
      p.throughput = xyz(t4);
      p.last_pdf = t4.w;
    }
    const float4 r4 = q.rad[slot];
    p.radiance = xyz(r4);
    p.rng = __float_as_uint(r4.w);
    const Scatter s = scatter_path(S, F, p, __float_as_uint(d4.w), h4.x, h4.y, __uint_as_float(h4.z),
                                   __uint_as_float(h4.w));
    const uint32_t ks = wave_append(shadow_count(counters, bounce), s.shadow);
    if (s.shadow) {
      q.sh_o[ks] = make_float4(s.shadow_ray.o.x, s.shadow_ray.o.y, s.shadow_ray.o.z, __uint_as_float(slot));
      q.sh_d[ks] = make_float4(s.shadow_ray.d.x, s.shadow_ray.d.y, s.shadow_ray.d.z, s.shadow_ray.tmax);
      q.sh_c[ks] = make_float4(s.delta.x, s.delta.y, s.delta.z, 0.0f);
    }
    const uint32_t kr = wave_append(ray_count(counters, bounce + 1), s.continues);
    if (s.continues) {
      q.thr[slot] = make_float4(p.throughput.x, p.throughput.y, p.throughput.z, p.last_pdf);
      q.rad[slot] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, __uint_as_float(p.rng));
      q.ray_o[kr] = make_float4(p.o.x, p.o.y, p.o.z, __uint_as_float(slot));
      q.ray_d[kr] = make_float4(p.d.x, p.d.y, p.d.z, kFloatMax);
    } else {
      q.rad[slot] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 1.0f);
    }
  }
}

// ShadowTransmittance (scene.slang:445-460) for the shadow queue of bounce
// b: an unblocked ray adds its light to its path's radiance.
__global__ void __launch_bounds__(kTraceBlock) k_shadow(SceneView S, Queues q, uint32_t* counters, uint2* spill, int spill_cap,
                                                        uint32_t bounce, unsigned long long* stats) {
  const uint32_t count = *shadow_count(counters, bounce);
  const Traverser<kStackRegs> tr{S.nodes, S.tris};
  TraversalStack<kStackRegs> st = make_stack<kStackRegs>(spill, spill_cap);
  for (uint32_t i = gtid(); i < count; i += gstride()) {
    const float4 o4 = q.sh_o[i], d4 = q.sh_d[i];
    const Ray r{xyz(o4), 0.0f, xyz(d4), d4.w};
    if (!tr.occluded(r, st)) {
      const uint32_t slot = __float_as_uint(o4.w);
      const float4 c4 = q.sh_c[i];
      float4 r4 = q.rad[slot];
      const V3 rad = xyz(r4) + xyz(c4);
      r4.x = rad.x;
      r4.y = rad.y;
      r4.z = rad.z;
      q.rad[slot] = r4;
    }
  }
  flush_stats(stats, 2, st);
}

// wavefront.slang:287-309: the paths still queued after the recorded rounds,
// each followed to its end in one invocation, shadow rays traced inline.
__global__ void __launch_bounds__(kTraceBlock) k_tail(SceneView S, FrameConst F, Queues q, uint32_t* counters,
                                                      uint2* spill, int spill_cap, uint32_t bounce, unsigned long long* stats) {
  const uint32_t count = *ray_count(counters, bounce);
  const Traverser<kStackRegs> tr{S.nodes, S.tris};
  TraversalStack<kStackRegs> st = make_stack<kStackRegs>(spill, spill_cap);
  const V3 env = V3{S.environment[0], S.environment[1], S.environment[2]};
  for (uint32_t i = gtid(); i < count; i += gstride()) {
    const float4 o4 = q.ray_o[i], d4 = q.ray_d[i];
    const uint32_t slot = __float_as_uint(o4.w);
    const float4 t4 = q.thr[slot], r4 = q.rad[slot];
    PathRegs p;
    p.o = xyz(o4);
    p.d = xyz(d4);
    p.throughput = xyz(t4);
    p.last_pdf = t4.w;
    p.radiance = xyz(r4);
    p.rng = __float_as_uint(r4.w);
    p.depth = bounce;
    while (true) {
      HitRecord h;
      if (!tr.closest(Ray{p.o, 0.0f, p.d, kFloatMax}, st, h)) {
        p.radiance = p.radiance + p.throughput * env;
        break;
      }
      const Scatter s = scatter_path(S, F, p, h.tri, h.instance, h.primitive, h.u, h.v);
      if (s.shadow) {
        atomicAdd(shadow_count(counters, p.depth - (s.continues ? 1u : 0u)), 1u);
        if (!tr.occluded(s.shadow_ray, st)) p.radiance = p.radiance + s.delta;
      }
      if (!s.continues) break;
      atomicAdd(ray_count(counters, p.depth), 1u);
    }
    q.rad[slot] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 1.0f);
  }
  flush_stats(stats, 3, st);
}

// path_trace.slang:106-139 (kResolvePass): adds each slot's sample to the
// pixel's accumulation and writes the mean, a missed sample counting as the
// background; pixels no sample has hit keep the clear colour.
__global__ void __launch_bounds__(256) k_resolve(FrameConst F, Queues q, uint32_t add_sample) {
  const uint32_t rw = F.rect[2], total = F.rect[2] * F.rect[3];
  for (uint32_t i = gtid(); i < total; i += gstride()) {
    const uint32_t x = F.rect[0] + i % rw, y = F.rect[1] + i / rw;
    const uint32_t slot = y * F.width + x;
    float4 sum = F.accumulated != 0 ? q.accum[slot] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (add_sample) {
      const float4 s = q.rad[slot];
      sum.x = sum.x + s.x;
      sum.y = sum.y + s.y;
      sum.z = sum.z + s.z;
      sum.w = sum.w + s.w;
      q.accum[slot] = sum;
    }
    const float count = static_cast<float>(F.accumulated + add_sample);
    if (sum.w == 0.0f || count == 0.0f) {
      q.color[slot] = make_float4(F.background[0], F.background[1], F.background[2], F.background[3]);
      continue;
    }
    const float miss = count - sum.w;
    q.color[slot] = make_float4((sum.x + miss * F.background[0]) / count, (sum.y + miss * F.background[1]) / count,
                                (sum.z + miss * F.background[2]) / count, (sum.w + miss * F.background[3]) / count);
  }
}

// path_trace.slang:69-97 (CameraPass): the pixel centre's closest hit as
// depth and barycentric weights or a surface diagnostic. Misses keep the
// clears (colour = background, depth = 1).
__global__ void __launch_bounds__(kTraceBlock) k_camera_pass(SceneView S, FrameConst F, Queues q, float* depth,
                                                             uint32_t write_color, uint2* spill, int spill_cap) {
  const uint32_t rw = F.rect[2], total = F.rect[2] * F.rect[3];
  const Traverser<kStackRegs> tr{S.nodes, S.tris};
  TraversalStack<kStackRegs> st = make_stack<kStackRegs>(spill, spill_cap);
  for (uint32_t i = gtid(); i < total; i += gstride()) {
    const uint32_t x = F.rect[0] + i % rw, y = F.rect[1] + i / rw;
    const uint32_t slot = y * F.width + x;
    const float fx = static_cast<float>(x) + 0.5f, fy = static_cast<float>(y) + 0.5f;
    const Ray r = camera_ray(F.clip_to_world, fx * F.ndc_from_pixel[0] + F.ndc_from_pixel[2],
                             fy * F.ndc_from_pixel[1] + F.ndc_from_pixel[3]);
    HitRecord h;
    if (!tr.closest(r, st, h)) continue;
    const V3 pos = r.o + r.d * h.t;
    float clip[4];
    mul_point(F.world_to_clip, pos.x, pos.y, pos.z, 1.0f, clip);
    if (depth) depth[slot] = clip[2] / clip[3];
    if (!write_color) continue;
    float4 c;
    if (F.output == 1u) {
      c = make_float4(1.0f - h.u - h.v, h.u, h.v, 1.0f);
    } else {
      const Surface sf = reconstruct_hit(S, h.tri, h.instance, h.u, h.v, r.d);
      if (F.output == 3u) c = make_float4(sf.base_color.x, sf.base_color.y, sf.base_color.z, 1.0f);
      else if (F.output == 4u) c = make_float4(sf.roughness, sf.metallic, 0.0f, 1.0f);
      else c = make_float4(sf.shading_normal.x, sf.shading_normal.y, sf.shading_normal.z, 1.0f);
    }
    q.color[slot] = c;
  }
}

// Trace / ShadowTransmittance over caller rays (8 floats in, 5 words or 1 out).
__global__ void __launch_bounds__(kTraceBlock) k_trace(SceneView S, const float4* rays, uint32_t* hits,
                                                       uint32_t count, uint32_t any_hit, uint2* spill, int spill_cap) {
  const Traverser<kStackRegs> tr{S.nodes, S.tris};
  TraversalStack<kStackRegs> st = make_stack<kStackRegs>(spill, spill_cap);
  for (uint32_t i = gtid(); i < count; i += gstride()) {
    const float4 o4 = rays[2 * i], d4 = rays[2 * i + 1];
    const Ray r{xyz(o4), o4.w, xyz(d4), d4.w};
    if (any_hit) {
      hits[i] = tr.occluded(r, st) ? 1u : 0u;
      continue;
    }
    HitRecord h;
    uint32_t* o = hits + 5ull * i;
    if (tr.closest(r, st, h)) {
      o[0] = h.instance;
      o[1] = h.primitive;
      o[2] = __float_as_uint(h.u);
      o[3] = __float_as_uint(h.v);
      o[4] = __float_as_uint(h.t);
    } else {
      o[0] = kMiss;
      o[1] = kMiss;
      o[2] = 0u;
      o[3] = 0u;
      o[4] = 0u;
    }
  }
}

__global__ void k_zero(uint32_t* p, uint32_t n) {
  for (uint32_t i = gtid(); i < n; i += gstride()) p[i] = 0u;
}

__global__ void k_fill4(float4* p, uint32_t n, float4 v) {
  for (uint32_t i = gtid(); i < n; i += gstride()) p[i] = v;
}

__global__ void k_fill1(float* p, uint32_t n, float v) {
  for (uint32_t i = gtid(); i < n; i += gstride()) p[i] = v;
}

}  // namespace

void launch_zero(uint32_t* p, uint32_t n, hipStream_t s) {
  k_zero<<<(n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024, 256, 0, s>>>(p, n);
}
void launch_fill4(float4* p, uint32_t n, float4 v, hipStream_t s) {
  k_fill4<<<(n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024, 256, 0, s>>>(p, n, v);
}
void launch_fill1(float* p, uint32_t n, float v, hipStream_t s) {
  k_fill1<<<(n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024, 256, 0, s>>>(p, n, v);
}
void launch_camera(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c) {
  k_camera<<<L.trace_blocks, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats);
}
void launch_intersect(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                      uint32_t bounce) {
  k_intersect<<<L.trace_blocks, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
}
void launch_shade(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                  uint32_t bounce) {
  k_shade<<<L.shade_blocks, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce);
}
void launch_shadow(const Launch& L, const SceneView& S, const Queues& q, uint32_t* c, uint32_t bounce) {
  k_shadow<<<L.trace_blocks, kTraceBlock, 0, L.stream>>>(S, q, c, L.spill, L.spill_cap, bounce, L.stats);
}
void launch_tail(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                 uint32_t bounce) {
  k_tail<<<L.trace_blocks, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
}
void launch_resolve(const Launch& L, const FrameConst& F, const Queues& q, uint32_t add_sample) {
  k_resolve<<<L.shade_blocks, 256, 0, L.stream>>>(F, q, add_sample);
}
void launch_camera_pass(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, float* depth,
                        uint32_t write_color) {
  k_camera_pass<<<L.trace_blocks, kTraceBlock, 0, L.stream>>>(S, F, q, depth, write_color, L.spill, L.spill_cap);
}
void launch_trace(const Launch& L, const SceneView& S, const float* rays, uint32_t* hits, uint32_t count,
                  uint32_t any_hit) {
  k_trace<<<L.trace_blocks, kTraceBlock, 0, L.stream>>>(S, reinterpret_cast<const float4*>(rays), hits, count,
                                                         any_hit, L.spill, L.spill_cap);
}

}  // namespace lotus
