// Launch interface of the wavefront kernels (kernels.hip).
#pragma once

#include "transport.hpp"

namespace lotus {

// Registers of the traversal stack before it spills, and block sizes.
constexpr int kStackRegs = 8;
constexpr int kTraceBlock = 256;
constexpr int kShadeBlock = 256;

// Per-target path state and queues, P = width * height entries each.
struct Queues {
  float4* thr;    // per slot: throughput, last pdf (< 0: unweighted)
  float4* rad;    // per slot: radiance, rng bits; once ended, sample and weight
  float4* ray_o;  // ray queue: origin, slot
  float4* ray_d;  //            direction, tmax
  float4* hit_o;  // hit queue: ray origin, slot
  float4* hit_d;  //            ray direction, triangle (BVH order)
  uint4* hit_h;   //            instance, primitive, u, v
  float4* sh_o;   // shadow queue: origin, slot
  float4* sh_d;   //               direction, distance
  float4* sh_c;   //               radiance delta if unblocked
  float4* accum;  // per pixel: radiance sum, hit samples
  float4* color;  // per pixel: mean (RGBA32F product)
};

struct Launch {
  hipStream_t stream;
  uint32_t trace_blocks;  // persistent grid of the traversal kernels
  uint32_t shade_blocks;
  uint2* spill;           // traversal stack spill, trace_blocks * kTraceBlock lanes
  int spill_cap;          // spill entries per lane
  unsigned long long* stats;  // node visits, triangle tests: camera, intersect, shadow, tail
};

void launch_zero(uint32_t* p, uint32_t n, hipStream_t s);
void launch_fill4(float4* p, uint32_t n, float4 v, hipStream_t s);
void launch_fill1(float* p, uint32_t n, float v, hipStream_t s);
void launch_camera(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c);
void launch_intersect(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                      uint32_t bounce);
void launch_shade(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                  uint32_t bounce);
void launch_shadow(const Launch& L, const SceneView& S, const Queues& q, uint32_t* c, uint32_t bounce);
void launch_tail(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                 uint32_t bounce);
void launch_resolve(const Launch& L, const FrameConst& F, const Queues& q, uint32_t add_sample);
void launch_camera_pass(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, float* depth,
                        uint32_t write_color);
void launch_trace(const Launch& L, const SceneView& S, const float* rays, uint32_t* hits, uint32_t count,
                  uint32_t any_hit);

}  // namespace lotus
