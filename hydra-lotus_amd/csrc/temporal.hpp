// Temporal frames on the device: the primary-surface product, history
// reprojection and validation, and SVGF (temporal accumulation, variance
// estimate, edge-aware a-trous filter, composite). Record layouts are the
// reference's (include/lotus/temporal.hpp:193-245), so its CPU oracles read
// them unchanged.
#pragma once

#include "kernels.hpp"

namespace lotus {

struct TemporalPixel {  // temporal.hpp:193-212, 64 bytes
  float4 normal_depth;   // unit geometric normal facing the camera, view depth
  float4 motion_depth;   // backward motion (px), expected previous depth, valid projection
  uint64_t instance_id, material_id;
  uint32_t flags;        // bit 0 hit, bit 1 rigid correspondence, bit 2 material unchanged
  uint32_t albedo;       // unorm8 RGB
  uint32_t shading_normal;  // octahedral snorm16 x, y
  uint32_t reserved;
};
static_assert(sizeof(TemporalPixel) == 64, "TemporalPixel is 64 bytes");

struct TemporalValidityPixel {  // temporal.hpp:214-220
  uint32_t valid, rejection, previous_index, reserved;
};

struct TemporalRadiancePixel {  // temporal.hpp:222-234
  float4 color;
  uint32_t sample_count;
  float first_moment, second_moment, variance;
};
static_assert(sizeof(TemporalRadiancePixel) == 32, "TemporalRadiancePixel is 32 bytes");

// temporal.slang:10-24 (TemporalConstants), the parts this build uses.
struct TemporalConst {
  float current_view[16];           // world -> view, column-major
  float previous_view[16];
  float previous_world_to_clip[16]; // Vulkan clip convention
  float previous_display[4];        // x, y, w, h of the previous frame's display window
  uint32_t previous_data[4];        // the previous frame's data window
  float validation[4];              // absolute / relative depth tolerance, normal cosine, max motion px
  uint32_t available;               // a compatible previous frame exists
  uint32_t width, height;
  uint32_t fresh_samples, cap, iterations;
  // TemporalHistoryRectification (include/lotus/temporal.hpp:35-57): 0 none,
  // 1 neighbourhood clamp, 2 anti-lag, 3 reflection weight; the world-space
  // camera positions the reflection weight compares.
  uint32_t rectification;
  float current_eye[3], previous_eye[3];
};
constexpr uint32_t kRectifyClamp = 1, kRectifyAntiLag = 2, kRectifyReflection = 3;

// Per-instance temporal identity (temporal.slang:32-40): previous placement
// relative to the current one, stable ids, flags (bit 1 rigid
// correspondence, bit 2 material unchanged). Null: instance + 1, material
// slot + 1, static placement.
struct TemporalInstances {
  const float* previous_from_current;  // 16 floats per instance
  const uint64_t* instance_id;
  const uint64_t* material_id;
  const uint32_t* flags;
};

void launch_temporal_surface(const Launch& L, const SceneView& S, const FrameConst& F, const TemporalConst& T,
                             const TemporalInstances& I, TemporalPixel* out);
void launch_temporal_reproject(const Launch& L, const TemporalConst& T, const TemporalPixel* current,
                               const TemporalPixel* previous, TemporalValidityPixel* validity);
void launch_temporal_settle(const Launch& L, const TemporalConst& T, const float4* prior,
                            const TemporalRadiancePixel* old_radiance, float4* color, uint32_t samples,
                            uint32_t max_samples);
void launch_temporal_accumulate(const Launch& L, const TemporalConst& T, const float4* batch,
                                const TemporalPixel* surface, const TemporalValidityPixel* validity,
                                const TemporalRadiancePixel* old_radiance, TemporalRadiancePixel* radiance);
// Variance words of `radiance`; with iterations > 0, also the filter's
// iteration-0 input and guides in `filter` (3 x width x height float4).
void launch_temporal_variance(const Launch& L, const TemporalConst& T, const TemporalPixel* surface,
                              TemporalRadiancePixel* radiance, float4* filter);
// Iterations 0 .. n-2 in place, then the composite (the last iteration,
// remodulated) into color; without iterations the accumulated colour.
void launch_temporal_filter(const Launch& L, const TemporalConst& T, const TemporalPixel* surface,
                            const TemporalRadiancePixel* radiance, float4* filter, float4* color);

}  // namespace lotus
