// Render-buffer writes of the renderer's products, on the device: the row
// flip and format conversion of HdLotusRenderBuffer::WriteColor, WriteDepth
// and WriteRowsFlippedLocked (adapters/hydra2/src/adapter.cpp:2544-2596,
// 2713-2731) and the id fill of WriteIds (:2733-2743).
//
// HBM-bound copies: one lane per pixel, grid-strided over rows so that a
// wave's loads and stores are both contiguous runs of one row.
#include "products.hpp"

namespace lotus {
namespace {

constexpr int kBlock = 256;

// lround(clamp(v, 0, 1) * 255) of the host code: a NaN passes std::clamp and
// lround turns it into the minimum long, whose low byte is 0.
__device__ __forceinline__ uint32_t to_unorm8(float v) {
  if (!(v == v)) return 0u;
  const float c = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  return static_cast<uint32_t>(roundf(c * 255.0f));
}

struct CopyPixel16 {
  __device__ static void run(const void* s, void* d, size_t si, size_t di) {
    static_cast<uint4*>(d)[di] = static_cast<const uint4*>(s)[si];
  }
};
struct CopyPixel4 {
  __device__ static void run(const void* s, void* d, size_t si, size_t di) {
    static_cast<uint32_t*>(d)[di] = static_cast<const uint32_t*>(s)[si];
  }
};
struct FloatToUnorm {
  __device__ static void run(const void* s, void* d, size_t si, size_t di) {
    const float4 v = static_cast<const float4*>(s)[si];
    static_cast<uint32_t*>(d)[di] = to_unorm8(v.x) | (to_unorm8(v.y) << 8) | (to_unorm8(v.z) << 16) |
                                    (to_unorm8(v.w) << 24);
  }
};
struct UnormToFloat {
  __device__ static void run(const void* s, void* d, size_t si, size_t di) {
    const uint32_t p = static_cast<const uint32_t*>(s)[si];
    static_cast<float4*>(d)[di] = make_float4(float(p & 0xFFu) / 255.0f, float((p >> 8) & 0xFFu) / 255.0f,
                                              float((p >> 16) & 0xFFu) / 255.0f, float(p >> 24) / 255.0f);
  }
};

// Grid: x covers a row in kBlock-lane chunks, y strides over rows.
template <class Op>
__global__ __launch_bounds__(kBlock) void k_write_flipped(const void* src, void* dst, uint32_t width, uint32_t height) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= width) return;
  for (uint32_t y = blockIdx.y; y < height; y += gridDim.y)
    Op::run(src, dst, size_t(y) * width + x, size_t(height - 1u - y) * width + x);
}

// HdLotusRenderBuffer::WriteTemporal (adapter.cpp:2598-2630) per pixel:
// lotus:motion is the surface record's motion_depth (RGBA32F); the others
// Int32: lotus:historyLength the radiance record's sample count (0 without
// radiance records), lotus:historyValidity and lotus:historyRejection the
// validity record's words. Records: TemporalPixel 64 B (motion_depth at 16),
// TemporalValidityPixel 16 B, TemporalRadiancePixel 32 B (sample_count at 16).
__global__ __launch_bounds__(kBlock) void k_temporal_aov(uint32_t aov, const uint4* surface, const uint4* validity,
                                                         const uint4* radiance, void* dst, uint32_t width,
                                                         uint32_t height) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= width) return;
  for (uint32_t y = blockIdx.y; y < height; y += gridDim.y) {
    const size_t i = size_t(y) * width + x, o = size_t(height - 1u - y) * width + x;
    if (aov == kAovMotion) {
      static_cast<uint4*>(dst)[o] = surface[4 * i + 1];
    } else {
      uint32_t v;
      if (aov == kAovHistoryLength) v = radiance ? radiance[2 * i + 1].x : 0u;
      else if (aov == kAovHistoryValidity) v = validity[i].x;
      else v = validity[i].y;
      static_cast<uint32_t*>(dst)[o] = v;
    }
  }
}

// HdLotusRenderBuffer::WriteRestir (adapter.cpp:2632-2672) per pixel: the
// resolved record is the spatial pass's, else the temporal pass's
// (RestirReusePixel, 64 B), else the initial RestirPixel (48 B), else an
// empty record (light 0xFFFFFFFF, zeros). lotus:restirWeight is
// {weight_sum, selected_target, normalization, proposal}, the normalization
// the reuse record's or (weight_sum / M) / selected_target;
// lotus:restirDirect the contribution; lotus:restirLight, lotus:restirCandidates,
// lotus:restirSupport (the reuse record's supporting candidates, else M)
// and lotus:restirAge (the reuse record's, else 0) Int32.
__global__ __launch_bounds__(kBlock) void k_restir_aov(uint32_t aov, const uint32_t* initial, const uint32_t* resolved,
                                                       void* dst, uint32_t width, uint32_t height) {
  const uint32_t x = blockIdx.x * kBlock + threadIdx.x;
  if (x >= width) return;
  for (uint32_t y = blockIdx.y; y < height; y += gridDim.y) {
    const size_t i = size_t(y) * width + x, o = size_t(height - 1u - y) * width + x;
    uint32_t r[12] = {0xFFFFFFFFu, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    uint32_t reuse[4] = {0u, 0u, 0xFFFFFFFFu, 0u};
    const uint32_t* rec = resolved ? resolved + 16 * i : (initial ? initial + 12 * i : nullptr);
    if (rec)
      for (int k = 0; k < 12; ++k) r[k] = rec[k];
    if (resolved)
      for (int k = 0; k < 4; ++k) reuse[k] = resolved[16 * i + 12 + k];
    const uint32_t m = r[3];
    const float weight_sum = __uint_as_float(r[1]), target = __uint_as_float(r[2]);
    if (aov == kAovRestirDirect) {
      static_cast<uint4*>(dst)[o] = make_uint4(r[8], r[9], r[10], r[11]);
    } else if (aov == kAovRestirWeight) {
      float normalization = 0.0f;
      if (resolved) normalization = __uint_as_float(reuse[0]);
      else if (m != 0u && target > 0.0f) normalization = (weight_sum / static_cast<float>(m)) / target;
      static_cast<uint4*>(dst)[o] = make_uint4(r[1], r[2], __float_as_uint(normalization), r[6]);
    } else {
      uint32_t v;
      if (aov == kAovRestirLight) v = r[0];
      else if (aov == kAovRestirCandidates) v = m;
      else if (aov == kAovRestirSupport) v = resolved ? reuse[1] : m;
      else v = resolved ? reuse[3] : 0u;
      static_cast<uint32_t*>(dst)[o] = v;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_ids(int32_t* dst, uint64_t count, int32_t value) {
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < count; i += uint64_t(gridDim.x) * kBlock)
    dst[i] = value;
}

}  // namespace

bool launch_write_rows_flipped(const void* src, uint32_t src_format, void* dst, uint32_t dst_format, uint32_t width,
                               uint32_t height, hipStream_t s) {
  // rows per block column: enough blocks to fill the chip at small widths
  const uint32_t cols = (width + kBlock - 1) / kBlock;
  const uint32_t rows = height < 4096u ? height : 4096u;
  const dim3 grid(cols, rows);
  if (src_format == dst_format) {
    if (src_format == kFloat32x4)
      k_write_flipped<CopyPixel16><<<grid, kBlock, 0, s>>>(src, dst, width, height);
    else
      k_write_flipped<CopyPixel4><<<grid, kBlock, 0, s>>>(src, dst, width, height);
  } else if (src_format == kFloat32x4 && dst_format == kUnorm8x4) {
    k_write_flipped<FloatToUnorm><<<grid, kBlock, 0, s>>>(src, dst, width, height);
  } else if (src_format == kUnorm8x4 && dst_format == kFloat32x4) {
    k_write_flipped<UnormToFloat><<<grid, kBlock, 0, s>>>(src, dst, width, height);
  } else {
    return false;
  }
  return true;
}

void launch_temporal_aov(uint32_t aov, const void* surface, const void* validity, const void* radiance, void* dst,
                         uint32_t width, uint32_t height, hipStream_t s) {
  const dim3 grid((width + kBlock - 1) / kBlock, height < 4096u ? height : 4096u);
  k_temporal_aov<<<grid, kBlock, 0, s>>>(aov, static_cast<const uint4*>(surface), static_cast<const uint4*>(validity),
                                         static_cast<const uint4*>(radiance), dst, width, height);
}

void launch_restir_aov(uint32_t aov, const void* initial, const void* resolved, void* dst, uint32_t width,
                       uint32_t height, hipStream_t s) {
  const dim3 grid((width + kBlock - 1) / kBlock, height < 4096u ? height : 4096u);
  k_restir_aov<<<grid, kBlock, 0, s>>>(aov, static_cast<const uint32_t*>(initial),
                                       static_cast<const uint32_t*>(resolved), dst, width, height);
}

void launch_fill_ids(int32_t* dst, uint64_t count, int32_t value, hipStream_t s) {
  uint64_t blocks = (count + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) return;
  k_fill_ids<<<static_cast<uint32_t>(blocks), kBlock, 0, s>>>(dst, count, value);
}

}  // namespace lotus
