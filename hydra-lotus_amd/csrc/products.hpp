// Render-buffer writes of the renderer's products (products.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lotus {

// Pixel formats of a product or a host render buffer (HdFormat's
// UNorm8Vec4, Float32Vec4, Float32 and Int32).
enum BufferFormat : uint32_t { kUnorm8x4 = 0, kFloat32x4 = 1, kFloat32 = 2, kInt32 = 3 };

// Bytes per pixel of a format, 0 for an unknown one.
inline uint32_t buffer_pixel_bytes(uint32_t f) {
  return f == kUnorm8x4 ? 4u : f == kFloat32x4 ? 16u : (f == kFloat32 || f == kInt32) ? 4u : 0u;
}

// Copies a top-down width x height product into a bottom-up buffer of the
// same size, converting RGBA32F to UNorm8 (clamp to [0, 1], times 255, round
// half away from zero) or UNorm8 to RGBA32F (v / 255); any other pair must
// be one format. Returns false for a pair it does not convert.
bool launch_write_rows_flipped(const void* src, uint32_t src_format, void* dst, uint32_t dst_format, uint32_t width,
                               uint32_t height, hipStream_t s);
// The named temporal and ReSTIR AOVs (adapter.cpp:256-285).
enum TemporalAov : uint32_t { kAovMotion = 0, kAovHistoryLength = 1, kAovHistoryValidity = 2, kAovHistoryRejection = 3 };
enum RestirAov : uint32_t {
  kAovRestirLight = 0,
  kAovRestirCandidates = 1,
  kAovRestirSupport = 2,
  kAovRestirAge = 3,
  kAovRestirWeight = 4,
  kAovRestirDirect = 5
};
inline uint32_t temporal_aov_format(uint32_t aov) { return aov == kAovMotion ? kFloat32x4 : kInt32; }
inline uint32_t restir_aov_format(uint32_t aov) {
  return aov == kAovRestirWeight || aov == kAovRestirDirect ? kFloat32x4 : kInt32;
}

// Writes a temporal AOV of width x height top-down records (TemporalPixel,
// TemporalValidityPixel, TemporalRadiancePixel or null) into a bottom-up
// buffer of the AOV's format.
void launch_temporal_aov(uint32_t aov, const void* surface, const void* validity, const void* radiance, void* dst,
                         uint32_t width, uint32_t height, hipStream_t s);
// Writes a ReSTIR AOV from the initial RestirPixel records and the resolved
// RestirReusePixel records (either may be null) into a bottom-up buffer.
void launch_restir_aov(uint32_t aov, const void* initial, const void* resolved, void* dst, uint32_t width,
                       uint32_t height, hipStream_t s);
// Fills `count` Int32 pixels with `value`.
void launch_fill_ids(int32_t* dst, uint64_t count, int32_t value, hipStream_t s);

}  // namespace lotus
