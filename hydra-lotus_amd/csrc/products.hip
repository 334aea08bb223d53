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

void launch_fill_ids(int32_t* dst, uint64_t count, int32_t value, hipStream_t s) {
  uint64_t blocks = (count + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) return;
  k_fill_ids<<<static_cast<uint32_t>(blocks), kBlock, 0, s>>>(dst, count, value);
}

}  // namespace lotus
