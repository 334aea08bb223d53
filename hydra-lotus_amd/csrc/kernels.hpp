// Launch interface of the wavefront kernels (kernels.hip).
#pragma once

#include <algorithm>
#include <type_traits>

#include "nrc.hpp"
#include "transport.hpp"

namespace lotus {

// LDS ring entries of a lane's traversal stack before it spills; idle lanes
// of a wave that trigger a refill; block sizes.
constexpr int kRing = 12;
constexpr int kRefill = 16;
constexpr int kTraceBlock = 256;
constexpr int kShadeBlock = 256;
// A traversal kernel's per-lane stacks (launched with kTraceBlock lanes).
using TraceStack = LdsStack<kRing, kTraceBlock>;

// Per-target path state and queues, P = width * height entries each.
struct Queues {
  float4* thr;    // per slot: throughput, last pdf (< 0: unweighted)
  float4* rad;    // per slot: radiance, rng bits; once ended, radiance and weight
  float4* nee;    // per slot: a shadow ray's light awaiting addition, flag
  float4* esc;    // per slot: the environment an escaping ray found; w: 1, or 2 when esc2 follows
  float4* esc2;   // per slot: the analytic lights that ray crossed (read only when esc.w is 2)
  float4* esc_o;  // escaped rays of a scene with lights, in escape order: origin + slot
  float4* esc_d;  //   and direction
  float4* ray_o;  // ray queue: origin, slot
  float4* ray_d;  //            direction, tmax
  float4* hit_o;  // hit queue: ray origin, slot
  float4* hit_d;  //            ray direction, triangle (BVH order)
  uint4* hit_h;   //            instance, primitive, u, v
  float4* sh_o;   // shadow queue: origin, slot
  float4* sh_d;   //               direction, distance
  float4* sh_c;   //               radiance delta if unblocked
  float4* sh_f;   //               with covered surfaces: f * radiance, MIS weight
  float4* sh_t;   //                 and the path's throughput, the light's pdf
  float4* accum;  // per pixel: radiance sum, hit samples
  float4* color;  // per pixel: mean (RGBA32F product)
  // Radiance cache frames.
  float4* cq;        // cache queries, 4 float4 each (position, direction, normal, material)
  float4* cq_thr;    // per query: path throughput, slot
  float4* crec;      // training vertices [bounce][path], 6 float4 each: query, throughput, radiance so far
  uint32_t* crec_n;  // per training path: vertices recorded
  float* samples;    // training samples, kSampleFloats each
  // ReSTIR DI records, 3 float4 per pixel (initial reservoirs).
  float4* restir;
};

// Training-path vertices recorded per path (deeper ones only contribute to
// the targets of these). Two: the cache answers queries at a path's second
// vertex, and in the equal-relMSE race on the Cornell box, recording 4 or 8
// vertices cost more training time (0.44 and 0.46 of plain's GPU time at 4
// frames) than it gained over 2 (0.36).
constexpr uint32_t kCacheRecordDepth = 2;
// Smallest throughput component a training vertex may divide by.
constexpr float kCacheMinThroughput = 1e-3f;

// Persistent grid of each kernel: workgroups filling every CU at its occupancy.
// Workgroups of each persistent launch: every CU at the kernel's own
// occupancy (a grid sized for another variant's occupancy would leave
// workgroups waiting for a CU, and a strided launch then ends a pass late).
struct GridShape {
  uint32_t camera, camera_dynamic, trace, trace_dynamic, tail, tail_cache, camera_pass, rays, shade, shade_parallel,
      resolve, paths;
  uint32_t paths_variant[3];  // the path megakernel's grids: plain, extras, covered surfaces
  uint32_t max_trace_lanes;  // the most traversal lanes any grid holds
};
GridShape grid_shape(uint32_t cus);

// BSDF sampling in a scene with several parallel lights: the shading lanes
// walk every parallel light's shadow ray themselves (the kernels' kParallel
// variants); with one, it is queued as next event estimation's ray is.
inline bool parallel_walks(const SceneView& S, const FrameConst& F) {
  return F.light_sampling == 0 && S.parallel_count > 1;
}

struct Launch {
  hipStream_t stream;
  GridShape grid;
  uint2* spill;               // traversal stack spill, max_trace_lanes lanes
  int spill_cap;              // spill entries per lane
  unsigned long long* stats;  // node visits, triangle tests: camera, extension, shadow, tail
  bool dynamic_camera;        // lanes hand over finished rays (else one ray per lane per pass)
  bool dynamic_trace;
  uint32_t camera_tile_shift = 3;  // camera tiles 2^shift pixels wide, 64 pixels each
};

// k_trace item kinds.
constexpr uint32_t kTraceExtension = 1u;
constexpr uint32_t kTraceShadow = 2u;
// Light sampling is on: escaping rays weigh the analytic lights they cross
// against it.
constexpr uint32_t kTraceWeigh = 4u;

void launch_zero(uint32_t* p, uint32_t n, hipStream_t s);
void launch_fill4(float4* p, uint32_t n, float4 v, hipStream_t s);
void launch_fill1(float* p, uint32_t n, float v, hipStream_t s);
void launch_camera(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c);
// An extension ray of the tail's first bounce whose path sent no shadow ray.
constexpr uint32_t kNoLink = 0xFFFFFFFFu;
// link: the last round before the tail, whose shadow rays the tail walks (a
// continuing path's extension ray carries its shadow ray's index in d.w).
void launch_shade(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                  uint32_t bounce, bool link = false);
// Extension rays of bounce `bounce` and/or shadow rays of bounce `bounce - 1`.
void launch_trace(const Launch& L, const SceneView& S, const Queues& q, uint32_t* c, uint32_t bounce,
                  uint32_t kinds);
void launch_tail(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                 uint32_t bounce);
// The pending environment and crossed analytic lights of bounce `bounce`'s
// escaped rays (scenes with lights; `weigh`: light sampling is on).
void launch_escape(const Launch& L, const SceneView& S, const Queues& q, uint32_t* c, uint32_t bounce, bool weigh);
void launch_resolve(const Launch& L, const FrameConst& F, const Queues& q, uint32_t add_sample);

// The path megakernel: every path of a frame from its camera ray to its end
// in one launch, for frames of BSDF or MIS light sampling without the
// radiance cache, reservoirs, glTF PBR or several parallel lights under BSDF
// sampling (paths_eligible); its results equal the wavefront's bit for bit.
// One view per launch, or a batch of views fed by one device work queue of
// kPathChunk-path chunks (a PathView table on the device, views in order).
constexpr uint32_t kPathChunk = 256;
constexpr uint32_t kXcds = 8;  // a fused launch's work-queue ranges: one per XCD (8 counters per launch)
struct PathView {
  SceneView S;
  FrameConst F;
  Queues q;
  uint32_t* counters;
  unsigned long long* stats;
  uint32_t first_chunk;  // the view's first chunk of the batch's queue
  uint32_t paths;        // its data window's paths, 8 x 4 tiles (path_chunks of them in chunks)
};
bool paths_eligible(const SceneView& S, const FrameConst& F);
uint32_t path_count(const FrameConst& F);   // paths of the data window (8 x 4 tiles, whole)
uint32_t path_chunks(const FrameConst& F);  // their chunks
void launch_paths(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c);
void launch_paths_views(const Launch& L, const PathView* views, uint32_t view_count, bool analytic, bool cov,
                        uint32_t* next_chunk, uint32_t total_chunks);
void launch_camera_pass(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, float* depth,
                        uint32_t write_color);
void launch_rays(const Launch& L, const SceneView& S, const float* rays, uint32_t* hits, uint32_t count,
                 uint32_t any_hit);
// ReSTIR DI's initial reservoirs at the first hits (queued by the camera
// kernel), with their visibility rays traced inline.
void launch_restir_initial(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c);
struct TemporalPixel;
struct TemporalValidityPixel;
// ReSTIR reuse passes of a temporal frame. Records: 4 float4 per pixel;
// surfaces: 8 float4 per pixel.
struct RestirReuseArgs {
  const float4* initial;
  const float4* surf;          // this frame's receivers
  const float4* prev_surf;     // the previous frame's receivers
  const float4* prev_hist;     // the previous frame's final records
  float4* temporal_out;
  float4* spatial_out;
  uint4* spatial_sources;      // the spatial pass's accepted neighbours, 2 uint4 per pixel
  const TemporalPixel* tsurf;  // this frame's pixel-centre surfaces
  const TemporalValidityPixel* validity;
  uint32_t history;            // the previous frame's records are usable
  uint32_t max_history;        // restir_max_history_candidates
  uint32_t radius;             // spatial radius in pixels
  uint32_t disk;               // 0: the four axis neighbours at the radius; 1: five drawn within it
};
void launch_restir_clear(const Launch& L, float4* restir, float4* surf, uint32_t pixels);
void launch_restir_temporal(const Launch& L, const SceneView& S, const FrameConst& F, const RestirReuseArgs& A,
                            uint32_t* c);
void launch_restir_spatial(const Launch& L, const SceneView& S, const FrameConst& F, const RestirReuseArgs& A,
                           uint32_t* c);
// Radiance cache frames: targets (L - R) / T of every recorded vertex of the
// training paths, appended to q.samples (cache_sample_count), at most `capacity`.
void launch_cache_samples(const Launch& L, const FrameConst& F, const Queues& q, uint32_t* c, uint32_t capacity);
// A frame's counters: per bounce b, kind 0 the ray-queue entries (camera rays
// at b = 0), kind 1 the hit-queue entries, kind 2 the shadow-queue entries;
// kind 3 of bounces 0, 1, 2 counts cache queries, cache samples and ReSTIR
// visibility rays. Every counter has a 256-byte line of its own: the queues
// are appended with one atomic per wave, and counters sharing a line would
// queue behind one another at one atomic unit.
#ifndef LOTUS_COUNTER_SPREAD
#define LOTUS_COUNTER_SPREAD 64u
#endif
constexpr uint32_t kCounterSpread = LOTUS_COUNTER_SPREAD;
__host__ __device__ constexpr uint32_t counter_word(uint32_t bounce, uint32_t kind) {
  return (4u * bounce + kind) * kCounterSpread;
}
__host__ __device__ inline uint32_t* cache_query_count(uint32_t* c) { return c + counter_word(0, 3); }
__host__ __device__ inline uint32_t* cache_sample_count(uint32_t* c) { return c + counter_word(1, 3); }
// ReSTIR visibility rays traced inline by the reservoir passes (not queued).
__host__ __device__ inline uint32_t* restir_ray_count(uint32_t* c) { return c + counter_word(2, 3); }
// Shadow rays the shading lanes walk themselves (several parallel lights
// under BSDF sampling): the second half of the camera-ray counter's line,
// whose only other atomics come from ray generation, before any shading.
__host__ __device__ inline uint32_t* parallel_ray_count(uint32_t* c) {
  return c + counter_word(0, 0) + kCounterSpread / 2;
}
// Extension rays of bounce b (b >= 1) that escaped a scene with analytic
// lights or an environment image: kind 3 of bounce b + 2.
__host__ __device__ inline uint32_t* escape_count(uint32_t* c, uint32_t bounce) { return c + counter_word(bounce + 2, 3); }
// Counter words a frame clears: 4 counters per bounce (and the escape counts
// one bounce further), at least the three above.
__host__ __device__ inline uint32_t frame_counter_words(uint32_t bounce_stride) {
  return counter_word((bounce_stride > 3u ? bounce_stride : 3u) + 1u, 0);
}

}  // namespace lotus
