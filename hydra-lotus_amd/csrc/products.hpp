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
// Fills `count` Int32 pixels with `value`.
void launch_fill_ids(int32_t* dst, uint64_t count, int32_t value, hipStream_t s);

}  // namespace lotus
