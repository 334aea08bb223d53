// Radiance cache network on gfx950: hash-grid encoding, MLP inference,
// backpropagation, weight-gradient GEMM and Adam.
//
// Every layer is computed transposed, D^T = W . X^T, with
// v_mfma_f32_16x16x32_f16 (and 16x16x16 for 16-wide operands): the weights
// are the A operand, a 16-query batch tile is the B operand, and the D tile a
// lane holds (rows 4g..4g+3 of column c, g = lane / 16, c = lane % 16) is, after
// ReLU and conversion, exactly the B operand of the next layer when B's 32-wide
// K block is laid out as slot (g, j) -> feature 32 kb + (j < 4 ? 4g + j : 16 +
// 4g + j - 4). The weights are stored in that slot order (the "operand
// images"), so activations never leave registers between layers, forwards and
// backwards. The four lanes of a query (c, c + 16, c + 32, c + 48) each encode
// the four grid levels whose features land in their slots, and the backward
// pass's input gradients land in the same lanes, which scatter them to the grid.
#include <cmath>
#include <cstring>
#include <vector>

#include "nrc.hpp"

namespace lotus {
namespace {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kFwdHalves = 16384;   // forward operand image
constexpr int kBwdHalves = 15360;   // backward operand image
constexpr int kActRows = 48 + 4 * 64;
constexpr int kDeltaRows = 4 * 64 + 16;
constexpr int kChunk = 512;         // samples per weight-gradient block
constexpr int kTileStride = 24;     // halves per row of a wave's transpose tile
constexpr int kInferTiles = 1;      // 16-query tiles per wave and pass
constexpr float kHalfMax = 65504.0f;

// Feature of K-block kb held in slot j of lane group g.
__host__ __device__ constexpr int kslot(int kb, int g, int j) { return 32 * kb + (j < 4 ? 4 * g + j : 12 + 4 * g + j); }

struct GridArgs {
  const h2* grid;
  float* grad;  // dense levels
  h2* grad16;   // hashed levels
  uint32_t res[kCacheLevels];
  uint32_t dense_mask;
  uint32_t table;  // entries per level, a power of two
};

__device__ __forceinline__ f4 mma32(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mma16(h4 a, h4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }

__device__ __forceinline__ _Float16 to_half(float x) { return static_cast<_Float16>(fminf(fmaxf(x, -kHalfMax), kHalfMax)); }
__device__ __forceinline__ _Float16 relu_half(float x) { return static_cast<_Float16>(fminf(fmaxf(x, 0.0f), kHalfMax)); }

__device__ __forceinline__ uint32_t corner_index(uint32_t x, uint32_t y, uint32_t z, uint32_t res, bool dense,
                                                 uint32_t table) {
  if (dense) {
    const uint32_t s = res + 1u;
    return x + y * s + z * s * s;
  }
  return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & (table - 1u);
}

struct Cell {
  uint32_t x, y, z;
  float wx, wy, wz;
};

// The cell of a level holding p in [0, 1]^3 and p's offsets inside it.
__device__ __forceinline__ Cell level_cell(float px, float py, float pz, uint32_t res) {
  const float r = static_cast<float>(res), top = static_cast<float>(res - 1u);
  const float ax = px * r, ay = py * r, az = pz * r;
  const float fx = fminf(floorf(ax), top), fy = fminf(floorf(ay), top), fz = fminf(floorf(az), top);
  return Cell{static_cast<uint32_t>(fx), static_cast<uint32_t>(fy), static_cast<uint32_t>(fz), ax - fx, ay - fy,
              az - fz};
}

__device__ __forceinline__ float corner_weight(const Cell& c, int k) {
  return (((k & 1) ? c.wx : 1.0f - c.wx) * ((k & 2) ? c.wy : 1.0f - c.wy)) * ((k & 4) ? c.wz : 1.0f - c.wz);
}

// Level i (0..3) of lane group g: 2g, 2g + 1, 8 + 2g, 9 + 2g. The resolution
// is picked with selects so the kernel argument array is indexed statically.
template <int I>
__device__ __forceinline__ void lane_level(const GridArgs& G, int g, int& level, uint32_t& res, bool& dense) {
  constexpr int l0 = (I < 2 ? 0 : 8) + (I & 1);
  level = l0 + 2 * g;
  res = g == 0 ? G.res[l0] : g == 1 ? G.res[l0 + 2] : g == 2 ? G.res[l0 + 4] : G.res[l0 + 6];
  dense = ((G.dense_mask >> level) & 1u) != 0;
}

template <int I>
__device__ __forceinline__ void encode_level(const GridArgs& G, float px, float py, float pz, int g, h8& out) {
  int level;
  uint32_t res;
  bool dense;
  lane_level<I>(G, g, level, res, dense);
  const Cell c = level_cell(px, py, pz, res);
  const h2* base = G.grid + static_cast<size_t>(level) * G.table;
  float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float w = corner_weight(c, k);
    const h2 v = base[corner_index(c.x + (k & 1), c.y + ((k >> 1) & 1), c.z + ((k >> 2) & 1), res, dense, G.table)];
    f0 += w * static_cast<float>(v[0]);
    f1 += w * static_cast<float>(v[1]);
  }
  out[2 * I] = to_half(f0);
  out[2 * I + 1] = to_half(f1);
}

__device__ __forceinline__ h8 encode(const GridArgs& G, float px, float py, float pz, int g) {
  h8 out;
  encode_level<0>(G, px, py, pz, g, out);
  encode_level<1>(G, px, py, pz, g, out);
  encode_level<2>(G, px, py, pz, g, out);
  encode_level<3>(G, px, py, pz, g, out);
  return out;
}

template <int I>
__device__ __forceinline__ void scatter_level(const GridArgs& G, float px, float py, float pz, int g, float d0,
                                              float d1) {
  if (d0 == 0.0f && d1 == 0.0f) return;
  int level;
  uint32_t res;
  bool dense;
  lane_level<I>(G, g, level, res, dense);
  const Cell c = level_cell(px, py, pz, res);
  if (dense) {
    // a dense cell's few entries gather thousands of samples: FP32 sums
    float* base = G.grad + 2ull * static_cast<size_t>(level) * G.table;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float w = corner_weight(c, k);
      const uint32_t e = corner_index(c.x + (k & 1), c.y + ((k >> 1) & 1), c.z + ((k >> 2) & 1), res, true, G.table);
      unsafeAtomicAdd(base + 2 * e, w * d0);
      unsafeAtomicAdd(base + 2 * e + 1, w * d1);
    }
    return;
  }
  // a hashed entry gathers a few samples: both features in one packed FP16
  // atomic, half the atomic traffic of FP32 pairs (the sums are not yet
  // divided by the batch size, so they stay far from FP16's lower range)
  __attribute__((address_space(1))) h2* base =
      (__attribute__((address_space(1))) h2*)(G.grad16 + static_cast<size_t>(level) * G.table);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float w = corner_weight(c, k);
    const uint32_t e = corner_index(c.x + (k & 1), c.y + ((k >> 1) & 1), c.z + ((k >> 2) & 1), res, false, G.table);
    (void)__builtin_amdgcn_global_atomic_fadd_v2f16(base + e, h2{to_half(w * d0), to_half(w * d1)});
  }
}

// A query's 16 floats; position clamped to the unit cube, the lane group's
// four extra inputs (direction, normal, roughness, base colour, metallic, 1).
struct QueryIn {
  float px, py, pz;
  h4 extra;
};

__device__ __forceinline__ QueryIn load_query(const float* q, bool valid, int g) {
  float v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = 0.0f;
  if (valid) {
    const float4* q4 = reinterpret_cast<const float4*>(q);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 t = q4[i];
      v[4 * i] = t.x;
      v[4 * i + 1] = t.y;
      v[4 * i + 2] = t.z;
      v[4 * i + 3] = t.w;
    }
  }
  QueryIn r;
  r.px = fminf(fmaxf(v[0], 0.0f), 1.0f);
  r.py = fminf(fmaxf(v[1], 0.0f), 1.0f);
  r.pz = fminf(fmaxf(v[2], 0.0f), 1.0f);
  const float one = valid ? 1.0f : 0.0f;
  const float e0 = g == 0 ? v[3] : g == 1 ? v[7] : g == 2 ? v[11] : 0.0f;
  const float e1 = g == 0 ? v[4] : g == 1 ? v[8] : g == 2 ? v[12] : 0.0f;
  const float e2 = g == 0 ? v[5] : g == 1 ? v[9] : g == 2 ? v[13] : 0.0f;
  const float e3 = g == 0 ? v[6] : g == 1 ? v[10] : g == 2 ? one : 0.0f;
  r.extra = h4{to_half(e0), to_half(e1), to_half(e2), to_half(e3)};
  return r;
}

__device__ __forceinline__ h8 pack_relu(f4 lo, f4 hi) {
  h8 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r[j] = relu_half(lo[j]);
    r[4 + j] = relu_half(hi[j]);
  }
  return r;
}

// Gradient through ReLU: D tiles lo / hi of a 32-wide K block, masked where
// the layer's output (held in the same slots) was zero.
__device__ __forceinline__ h8 pack_masked(f4 lo, f4 hi, h8 act) {
  h8 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r[j] = act[j] > static_cast<_Float16>(0.0f) ? to_half(lo[j]) : static_cast<_Float16>(0.0f);
    r[4 + j] = act[4 + j] > static_cast<_Float16>(0.0f) ? to_half(hi[j]) : static_cast<_Float16>(0.0f);
  }
  return r;
}

// Operand image offsets, in 16-byte (h8) units; 16-wide K fragments in 8-byte
// (h4) units. Forward: layer 0 grid block [m], layer 0 extras [m] (h4),
// hidden layer l in 1..3 [m][kb], output [kb]. Backward: W4^T [m] (h4), then
// W3^T, W2^T, W1^T [m][kb], then the grid rows of W0^T [m][kb].
__device__ __forceinline__ int fwd_l0(int m, int lane) { return 64 * m + lane; }
__device__ __forceinline__ int fwd_l0e(int m, int lane) { return 512 + 64 * m + lane; }
__device__ __forceinline__ int fwd_hid(int l, int m, int kb, int lane) { return 384 + 512 * (l - 1) + 128 * m + 64 * kb + lane; }
__device__ __forceinline__ int fwd_out(int kb, int lane) { return 1920 + 64 * kb + lane; }
__device__ __forceinline__ int bwd_t4(int m, int lane) { return 64 * m + lane; }
__device__ __forceinline__ int bwd_hid(int l, int m, int kb, int lane) { return 128 + 512 * (3 - l) + 128 * m + 64 * kb + lane; }
__device__ __forceinline__ int bwd_t0(int m, int kb, int lane) { return 1664 + 128 * m + 64 * kb + lane; }

// The network on NB batch tiles: act[l] = the input of weight layer l + 1.
template <int NB>
__device__ __forceinline__ void forward(const h8* wf, const h8 (&x0)[NB], const h4 (&xe)[NB], h8 (&act)[4][2][NB],
                                        f4 (&out)[NB]) {
  const int lane = __lane_id();
  const h4* wf4 = reinterpret_cast<const h4*>(wf);
  const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  {
    f4 acc[4][NB];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const h8 a = wf[fwd_l0(m, lane)];
      const h4 e = wf4[fwd_l0e(m, lane)];
#pragma unroll
      for (int t = 0; t < NB; ++t) acc[m][t] = mma16(e, xe[t], mma32(a, x0[t], zero));
    }
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      act[0][0][t] = pack_relu(acc[0][t], acc[1][t]);
      act[0][1][t] = pack_relu(acc[2][t], acc[3][t]);
    }
  }
#pragma unroll
  for (int l = 1; l < 4; ++l) {
    f4 acc[4][NB];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int t = 0; t < NB; ++t) acc[m][t] = zero;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        const h8 a = wf[fwd_hid(l, m, kb, lane)];
#pragma unroll
        for (int t = 0; t < NB; ++t) acc[m][t] = mma32(a, act[l - 1][kb][t], acc[m][t]);
      }
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      act[l][0][t] = pack_relu(acc[0][t], acc[1][t]);
      act[l][1][t] = pack_relu(acc[2][t], acc[3][t]);
    }
  }
#pragma unroll
  for (int t = 0; t < NB; ++t) out[t] = zero;
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    const h8 a = wf[fwd_out(kb, lane)];
#pragma unroll
    for (int t = 0; t < NB; ++t) out[t] = mma32(a, act[3][kb][t], out[t]);
  }
}

__device__ __forceinline__ uint32_t sample_count(uint32_t n_host, const uint32_t* n_dev, uint32_t first,
                                                 uint32_t stride, uint32_t cap) {
  const uint32_t total = n_dev ? (*n_dev < n_host ? *n_dev : n_host) : n_host;
  const uint32_t n = total > first ? (total - first + stride - 1u) / stride : 0u;
  return n < cap ? n : cap;
}

__device__ __forceinline__ void copy_image(h8* dst, const _Float16* src, int halves) {
  const h8* s = reinterpret_cast<const h8*>(src);
  for (int i = threadIdx.x; i < halves / 8; i += blockDim.x) dst[i] = s[i];
}

__global__ void __launch_bounds__(256) k_cache_infer(GridArgs G, const _Float16* wimg, const float* queries, uint32_t n,
                                                     const uint32_t* n_dev, float* out, const float4* thr,
                                                     float4* rad) {
  __shared__ h8 wf[kFwdHalves / 8];
  copy_image(wf, wimg, kFwdHalves);
  __syncthreads();
  const uint32_t count = n_dev ? *n_dev : n;
  const int lane = __lane_id(), g = lane >> 4, c = lane & 15;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  constexpr uint32_t kPass = 16 * kInferTiles;
  for (uint32_t base = wave * kPass; base < count; base += waves * kPass) {
    h8 x0[kInferTiles];
    h4 xe[kInferTiles];
#pragma unroll
    for (int t = 0; t < kInferTiles; ++t) {
      const uint32_t q = base + 16 * t + c;
      const QueryIn in = load_query(queries + static_cast<size_t>(q) * kQueryFloats, q < count, g);
      x0[t] = encode(G, in.px, in.py, in.pz, g);
      xe[t] = in.extra;
    }
    h8 act[4][2][kInferTiles];
    f4 o[kInferTiles];
    forward<kInferTiles>(wf, x0, xe, act, o);
    if (g == 0) {
#pragma unroll
      for (int t = 0; t < kInferTiles; ++t) {
        const uint32_t q = base + 16 * t + c;
        if (q >= count) continue;
        if (rad) {
          const float4 t4 = thr[q];
          const uint32_t slot = __float_as_uint(t4.w);
          float4 r = rad[slot];
          r.x = r.x + t4.x * fmaxf(o[t][0], 0.0f);
          r.y = r.y + t4.y * fmaxf(o[t][1], 0.0f);
          r.z = r.z + t4.z * fmaxf(o[t][2], 0.0f);
          rad[slot] = r;
        } else {
          out[3 * static_cast<size_t>(q)] = o[t][0];
          out[3 * static_cast<size_t>(q) + 1] = o[t][1];
          out[3 * static_cast<size_t>(q) + 2] = o[t][2];
        }
      }
    }
  }
}

// Transposes a wave's 16-sample operand rows through LDS into [row][sample]
// storage. put*: lane (g, c) writes its slots of sample c; flush_rows: lane r
// stores row r's 16 samples (two 16-byte stores).
__device__ __forceinline__ void put8(_Float16* tile, h8 v, int kb, int g, int c) {
#pragma unroll
  for (int j = 0; j < 8; ++j) tile[kslot(kb, g, j) * kTileStride + c] = v[j];
}
__device__ __forceinline__ void put4(_Float16* tile, h4 v, int row0, int g, int c) {
#pragma unroll
  for (int j = 0; j < 4; ++j) tile[(row0 + 4 * g + j) * kTileStride + c] = v[j];
}
__device__ __forceinline__ void flush_rows(const _Float16* tile, int rows, _Float16* dst, uint32_t ld, uint32_t base) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int r = __lane_id();
  if (r < rows) {
    const uint4* s = reinterpret_cast<const uint4*>(tile + r * kTileStride);
    uint4* d = reinterpret_cast<uint4*>(dst + static_cast<size_t>(r) * ld + base);
    const uint4 a = s[0], b = s[1];
    d[0] = a;
    d[1] = b;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward and backward pass of 16-sample tiles: stores every layer's input
// (acts) and output gradient (deltas) for the weight-gradient GEMM, scatters
// the input gradients into the grid, and sums the loss. Sample i of the batch
// is samples[first + stride * i]; tiles past the count (padded to 32) carry
// zero inputs and zero gradients.
__global__ void __launch_bounds__(256) k_cache_backprop(GridArgs G, const _Float16* wfimg, const _Float16* wbimg,
                                                        const float* samples, uint32_t n_host, const uint32_t* n_dev,
                                                        uint32_t first, uint32_t stride, uint32_t cap,
                                                        _Float16* acts, _Float16* deltas, uint32_t ld, float* loss) {
  __shared__ h8 wf[kFwdHalves / 8];
  __shared__ h8 wb[kBwdHalves / 8];
  __shared__ __align__(16) _Float16 tiles[4][64 * kTileStride];
  copy_image(wf, wfimg, kFwdHalves);
  copy_image(wb, wbimg, kBwdHalves);
  __syncthreads();
  const uint32_t n = sample_count(n_host, n_dev, first, stride, cap);
  const uint32_t npad = (n + 31u) & ~31u;
  const int lane = __lane_id(), g = lane >> 4, c = lane & 15;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  _Float16* tile = tiles[threadIdx.x >> 6];
  const h4* wb4 = reinterpret_cast<const h4*>(wb);
  const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  float lsum = 0.0f;
  for (uint32_t tix = wave; tix < npad / 16u; tix += waves) {
    const uint32_t base = tix * 16u, s = base + c;
    const bool valid = s < n;
    const float* rec = samples + static_cast<size_t>(first + stride * (valid ? s : 0u)) * kSampleFloats;
    const QueryIn in = load_query(rec, valid, g);
    h8 x0[1] = {encode(G, in.px, in.py, in.pz, g)};
    h4 xe[1] = {in.extra};
    put8(tile, x0[0], 0, g, c);
    put4(tile, xe[0], 32, g, c);
    flush_rows(tile, 48, acts, ld, base);
    h8 act[4][2][1];
    f4 out[1];
    forward<1>(wf, x0, xe, act, out);
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      put8(tile, act[l][0][0], 0, g, c);
      put8(tile, act[l][1][0], 1, g, c);
      flush_rows(tile, 64, acts + static_cast<size_t>(48 + 64 * l) * ld, ld, base);
    }
    // relative L2: d/dp (p - y)^2 / (sg(p)^2 + 0.01), mean over the 3 channels
    f4 d5 = zero;
    if (g == 0 && valid) {
      const float4 y4 = reinterpret_cast<const float4*>(rec)[4];
      const float y[3] = {y4.x, y4.y, y4.z};
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const float p = out[0][j], den = p * p + 0.01f, e = p - y[j];
        const float d = 2.0f * e / den * (1.0f / 3.0f);
        d5[j] = fminf(fmaxf(d, -kCacheDeltaClamp), kCacheDeltaClamp);
        lsum += e * e / den;
      }
    }
    const h4 dz5 = h4{to_half(d5[0]), to_half(d5[1]), to_half(d5[2]), to_half(d5[3])};
    put4(tile, dz5, 0, g, c);
    flush_rows(tile, 16, deltas + static_cast<size_t>(256) * ld, ld, base);
    h8 dz[2];
    {
      f4 acc[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) acc[m] = mma16(wb4[bwd_t4(m, lane)], dz5, zero);
      dz[0] = pack_masked(acc[0], acc[1], act[3][0][0]);
      dz[1] = pack_masked(acc[2], acc[3], act[3][1][0]);
    }
    put8(tile, dz[0], 0, g, c);
    put8(tile, dz[1], 1, g, c);
    flush_rows(tile, 64, deltas + static_cast<size_t>(192) * ld, ld, base);
#pragma unroll
    for (int l = 3; l >= 1; --l) {
      f4 acc[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        acc[m] = zero;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) acc[m] = mma32(wb[bwd_hid(l, m, kb, lane)], dz[kb], acc[m]);
      }
      dz[0] = pack_masked(acc[0], acc[1], act[l - 1][0][0]);
      dz[1] = pack_masked(acc[2], acc[3], act[l - 1][1][0]);
      put8(tile, dz[0], 0, g, c);
      put8(tile, dz[1], 1, g, c);
      flush_rows(tile, 64, deltas + static_cast<size_t>(64 * (l - 1)) * ld, ld, base);
    }
    f4 gacc[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      gacc[m] = zero;
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) gacc[m] = mma32(wb[bwd_t0(m, kb, lane)], dz[kb], gacc[m]);
    }
    if (valid) {
      scatter_level<0>(G, in.px, in.py, in.pz, g, gacc[0][0], gacc[0][1]);
      scatter_level<1>(G, in.px, in.py, in.pz, g, gacc[0][2], gacc[0][3]);
      scatter_level<2>(G, in.px, in.py, in.pz, g, gacc[1][0], gacc[1][1]);
      scatter_level<3>(G, in.px, in.py, in.pz, g, gacc[1][2], gacc[1][3]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (lane == 0 && lsum != 0.0f) atomicAdd(loss, lsum);
}

// dW[o][i] over samples [k0, k1) of one layer: A = deltas rows (o), B = acts
// rows (i), both read as 8 consecutive samples per lane.
template <int MO, int MI>
__device__ __forceinline__ void wgrad_layer(const _Float16* D, const _Float16* X, uint32_t ld, uint32_t k0, uint32_t k1,
                                            float* part, int in_width) {
  const int lane = __lane_id(), g = lane >> 4, r = lane & 15;
  f4 acc[MO][MI];
#pragma unroll
  for (int mo = 0; mo < MO; ++mo)
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) acc[mo][mi] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t k = k0; k < k1; k += 32u) {
    h8 a[MO], b[MI];
#pragma unroll
    for (int mo = 0; mo < MO; ++mo)
      a[mo] = *reinterpret_cast<const h8*>(D + static_cast<size_t>(16 * mo + r) * ld + k + 8 * g);
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
      b[mi] = *reinterpret_cast<const h8*>(X + static_cast<size_t>(16 * mi + r) * ld + k + 8 * g);
#pragma unroll
    for (int mo = 0; mo < MO; ++mo)
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) acc[mo][mi] = mma32(a[mo], b[mi], acc[mo][mi]);
  }
#pragma unroll
  for (int mo = 0; mo < MO; ++mo)
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
      for (int j = 0; j < 4; ++j) part[(16 * mo + 4 * g + j) * in_width + 16 * mi + r] = acc[mo][mi][j];
}

// One block per sample chunk; wave 0 takes layers 0 and 4, waves 1..3 the
// hidden layers. Partial sums go to part[block][kCacheParams].
__global__ void __launch_bounds__(256) k_cache_wgrad(const _Float16* acts, const _Float16* deltas, uint32_t ld,
                                                     uint32_t n_host, const uint32_t* n_dev, uint32_t first,
                                                     uint32_t stride, uint32_t cap, float* part) {
  const uint32_t n = sample_count(n_host, n_dev, first, stride, cap);
  const uint32_t npad = (n + 31u) & ~31u;
  const uint32_t k0 = blockIdx.x * kChunk;
  const uint32_t k1 = k0 < npad ? (npad - k0 < static_cast<uint32_t>(kChunk) ? npad : k0 + kChunk) : k0;
  float* p = part + static_cast<size_t>(blockIdx.x) * kCacheParams;
  const int wave = threadIdx.x >> 6;
  if (wave == 0) {
    wgrad_layer<4, 3>(deltas, acts, ld, k0, k1, p, 48);
    wgrad_layer<1, 4>(deltas + static_cast<size_t>(256) * ld, acts + static_cast<size_t>(240) * ld, ld, k0, k1,
                      p + kCacheWeightOffset[4], 64);
  } else {
    wgrad_layer<4, 4>(deltas + static_cast<size_t>(64 * wave) * ld, acts + static_cast<size_t>(48 + 64 * (wave - 1)) * ld,
                      ld, k0, k1, p + kCacheWeightOffset[wave], 64);
  }
}

__global__ void __launch_bounds__(256) k_cache_wsum(const float* part, uint32_t chunks, float* wgrad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= static_cast<uint32_t>(kCacheParams)) return;
  float s = 0.0f;
  for (uint32_t b = 0; b < chunks; ++b) s += part[static_cast<size_t>(b) * kCacheParams + i];
  wgrad[i] = s;
}

struct AdamArgs {
  float lr, beta1, beta2, epsilon;
  float c1, c2;  // 1 / (1 - beta^t)
};

__device__ __forceinline__ void adam_update(float& p, float g, float& m, float& v, const AdamArgs& a) {
  m = a.beta1 * m + (1.0f - a.beta1) * g;
  v = a.beta2 * v + (1.0f - a.beta2) * (g * g);
  p = p - a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.epsilon);
}

__global__ void __launch_bounds__(256) k_cache_adam_w(float* w, const float* grad, float* m, float* v, AdamArgs a,
                                                      uint32_t n_host, const uint32_t* n_dev, uint32_t first,
                                                      uint32_t stride, uint32_t cap) {
  const uint32_t n = sample_count(n_host, n_dev, first, stride, cap);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (n == 0 || i >= static_cast<uint32_t>(kCacheParams)) return;
  float p = w[i], mi = m[i], vi = v[i];
  adam_update(p, grad[i] / static_cast<float>(n), mi, vi, a);
  w[i] = p;
  m[i] = mi;
  v[i] = vi;
}

// Sparse Adam of the grid: entries no sample touched (zero gradient) keep
// their moments; touched gradients are cleared for the next batch.
__global__ void __launch_bounds__(256) k_cache_adam_grid(float* p, _Float16* p16, float* grad, _Float16* grad16,
                                                         uint32_t dense_mask, uint32_t level_shift, float* m, float* v,
                                                         uint64_t count, AdamArgs a, uint32_t n_host,
                                                         const uint32_t* n_dev, uint32_t first, uint32_t stride,
                                                         uint32_t cap) {
  const uint32_t n = sample_count(n_host, n_dev, first, stride, cap);
  const float inv = n ? 1.0f / static_cast<float>(n) : 0.0f;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const bool dense = ((dense_mask >> static_cast<uint32_t>(i >> level_shift)) & 1u) != 0;
    const float g = dense ? grad[i] : static_cast<float>(grad16[i]);
    if (g == 0.0f) continue;
    if (dense) grad[i] = 0.0f;
    else grad16[i] = static_cast<_Float16>(0.0f);
    if (n == 0) continue;
    float x = p[i], mi = m[i], vi = v[i];
    adam_update(x, g * inv, mi, vi, a);
    p[i] = x;
    m[i] = mi;
    v[i] = vi;
    p16[i] = to_half(x);
  }
}

// Operand images from the FP32 master weights.
__global__ void __launch_bounds__(256) k_cache_pack(const float* w, _Float16* wf, _Float16* wb) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < kFwdHalves) {
    int row, col, layer;
    if (idx < 2048) {
      const int m = idx / 512, lane = (idx / 8) % 64, j = idx % 8;
      layer = 0, row = 16 * m + lane % 16, col = kslot(0, lane / 16, j);
    } else if (idx < 3072) {
      const int t = idx - 2048, m = t / 256, lane = (t / 4) % 64, j = t % 4;
      layer = 0, row = 16 * m + lane % 16, col = 32 + 4 * (lane / 16) + j;
    } else if (idx < 15360) {
      const int t = idx - 3072, u = t % 4096, m = u / 1024, kb = (u / 512) % 2, lane = (u / 8) % 64, j = u % 8;
      layer = 1 + t / 4096, row = 16 * m + lane % 16, col = kslot(kb, lane / 16, j);
    } else {
      const int t = idx - 15360, kb = t / 512, lane = (t / 8) % 64, j = t % 8;
      layer = 4, row = lane % 16, col = kslot(kb, lane / 16, j);
    }
    const int in = layer == 0 ? 48 : 64;
    wf[idx] = to_half(w[kCacheWeightOffset[layer] + row * in + col]);
    return;
  }
  const int b = idx - kFwdHalves;
  if (b >= kBwdHalves) return;
  int r, k, layer;  // element W_layer[k][r] (transposed operand: row r = input feature)
  if (b < 1024) {
    const int m = b / 256, lane = (b / 4) % 64, j = b % 4;
    layer = 4, r = 16 * m + lane % 16, k = 4 * (lane / 16) + j;
  } else if (b < 13312) {
    const int t = b - 1024, u = t % 4096, m = u / 1024, kb = (u / 512) % 2, lane = (u / 8) % 64, j = u % 8;
    layer = 3 - t / 4096, r = 16 * m + lane % 16, k = kslot(kb, lane / 16, j);
  } else {
    const int t = b - 13312, m = t / 1024, kb = (t / 512) % 2, lane = (t / 8) % 64, j = t % 8;
    layer = 0, r = 16 * m + lane % 16, k = kslot(kb, lane / 16, j);
  }
  const int in = layer == 0 ? 48 : 64;
  wb[b] = to_half(w[kCacheWeightOffset[layer] + k * in + r]);
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__global__ void __launch_bounds__(256) k_cache_init_grid(float* p, _Float16* p16, float* m, float* v, float* g,
                                                         uint64_t count, uint32_t seed) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t h = mix32(static_cast<uint32_t>(i) ^ mix32(seed + static_cast<uint32_t>(i >> 32) * 0x9e3779b9u));
    const float x = (static_cast<float>(h >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f) * 1e-4f;
    p[i] = x;
    p16[i] = to_half(x);
    m[i] = 0.0f;
    v[i] = 0.0f;
    g[i] = 0.0f;
  }
}

__global__ void __launch_bounds__(256) k_cache_half(const float* p, _Float16* p16, uint64_t count) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p16[i] = to_half(p[i]);
}

__global__ void __launch_bounds__(256) k_cache_zero(float* p, uint64_t count) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = 0.0f;
}

__global__ void __launch_bounds__(256) k_cache_zero16(_Float16* p, uint64_t count) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = static_cast<_Float16>(0.0f);
}

// The hashed levels' FP16 sums copied into the FP32 gradient array, which
// then holds every level (for reading the gradients out).
__global__ void __launch_bounds__(256) k_cache_grad_merge(float* grad, const _Float16* grad16, uint32_t dense_mask,
                                                          uint32_t level_shift, uint64_t count) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    if (((dense_mask >> static_cast<uint32_t>(i >> level_shift)) & 1u) == 0) grad[i] = static_cast<float>(grad16[i]);
}

GridArgs grid_args(const CacheState& c) {
  GridArgs G;
  G.grid = reinterpret_cast<const h2*>(c.grid16);
  G.grad = c.ggrad;
  G.grad16 = reinterpret_cast<h2*>(c.ggrad16);
  for (int l = 0; l < kCacheLevels; ++l) G.res[l] = c.resolution[l];
  G.dense_mask = c.dense_mask;
  G.table = 1u << c.config.log2_table;
  return G;
}

uint32_t cu_count() {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return cus > 0 ? static_cast<uint32_t>(cus) : 1u;
}

uint32_t stream_blocks(uint64_t count) {
  const uint64_t b = (count + 255) / 256;
  return static_cast<uint32_t>(b < 8192 ? (b ? b : 1) : 8192);
}

AdamArgs adam_args(const CacheState& c) {
  AdamArgs a;
  a.lr = c.config.learning_rate;
  a.beta1 = c.config.beta1;
  a.beta2 = c.config.beta2;
  a.epsilon = c.config.epsilon;
  a.c1 = static_cast<float>(1.0 / (1.0 - std::pow(static_cast<double>(c.config.beta1), c.step)));
  a.c2 = static_cast<float>(1.0 / (1.0 - std::pow(static_cast<double>(c.config.beta2), c.step)));
  return a;
}

template <typename T>
hipError_t alloc(T*& p, size_t count) {
  p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(&p), (count ? count : 1) * sizeof(T));
}

}  // namespace

void cache_free(CacheState& c) {
  for (float** p : {&c.w32, &c.wgrad, &c.wm, &c.wv, &c.grid32, &c.ggrad, &c.gm, &c.gv, &c.partials, &c.loss})
    if (*p) (void)hipFree(*p), *p = nullptr;
  for (_Float16** p : {&c.wf16, &c.wb16, &c.grid16, &c.ggrad16, &c.acts, &c.deltas})
    if (*p) (void)hipFree(*p), *p = nullptr;
  if (c.count) (void)hipFree(c.count), c.count = nullptr;
  c.ready = false;
}

bool cache_init(CacheState& c, const CacheConfig& cfg, hipStream_t s, std::string& err) {
  cache_free(c);
  if (cfg.log2_table < 10 || cfg.log2_table > 22) {
    err = "log2_table must lie in [10, 22]";
    return false;
  }
  if (cfg.min_resolution < 2 || cfg.max_resolution < cfg.min_resolution || cfg.max_resolution > (1u << 20)) {
    err = "resolutions must satisfy 2 <= min <= max <= 2^20";
    return false;
  }
  if (cfg.batch_capacity == 0 || cfg.batch_capacity > (1u << 22)) {
    err = "batch_capacity must lie in [1, 2^22]";
    return false;
  }
  if (!(cfg.learning_rate > 0.0f) || !(cfg.beta1 >= 0.0f && cfg.beta1 < 1.0f) || !(cfg.beta2 >= 0.0f && cfg.beta2 < 1.0f) ||
      !(cfg.epsilon >= 0.0f)) {
    err = "invalid Adam parameters";
    return false;
  }
  c.config = cfg;
  const double table = std::ldexp(1.0, static_cast<int>(cfg.log2_table));
  const double growth = std::exp((std::log(static_cast<double>(cfg.max_resolution)) -
                                  std::log(static_cast<double>(cfg.min_resolution))) / (kCacheLevels - 1));
  c.dense_mask = 0;
  for (int l = 0; l < kCacheLevels; ++l) {
    double r = std::floor(cfg.min_resolution * std::pow(growth, l) + 1e-9);
    if (r > cfg.max_resolution) r = cfg.max_resolution;
    c.resolution[l] = static_cast<uint32_t>(r);
    const double cells = std::pow(r + 1.0, 3.0);
    if (cells <= table) c.dense_mask |= 1u << l;
  }
  c.grid_entries = static_cast<uint64_t>(kCacheLevels) * (1ull << cfg.log2_table) * kCacheFeatures;
  c.ld = (cfg.batch_capacity + 31u) & ~31u;
  c.chunks = (c.ld + kChunk - 1) / kChunk;
  c.step = 0;
  hipError_t e = alloc(c.w32, kCacheParams);
  for (float** p : {&c.wgrad, &c.wm, &c.wv})
    if (e == hipSuccess) e = alloc(*p, kCacheParams);
  if (e == hipSuccess) e = alloc(c.wf16, kFwdHalves);
  if (e == hipSuccess) e = alloc(c.wb16, kBwdHalves);
  for (float** p : {&c.grid32, &c.ggrad, &c.gm, &c.gv})
    if (e == hipSuccess) e = alloc(*p, c.grid_entries);
  if (e == hipSuccess) e = alloc(c.grid16, c.grid_entries);
  if (e == hipSuccess) e = alloc(c.ggrad16, c.grid_entries);
  if (e == hipSuccess) e = alloc(c.acts, static_cast<size_t>(kActRows) * c.ld);
  if (e == hipSuccess) e = alloc(c.deltas, static_cast<size_t>(kDeltaRows) * c.ld);
  if (e == hipSuccess) e = alloc(c.partials, static_cast<size_t>(c.chunks) * kCacheParams);
  if (e == hipSuccess) e = alloc(c.loss, 1);
  if (e == hipSuccess) e = alloc(c.count, 1);
  if (e != hipSuccess) {
    err = std::string("radiance cache allocation: ") + hipGetErrorString(e);
    cache_free(c);
    return false;
  }
  // Xavier-uniform hidden weights from a counter hash; the output layer starts
  // at zero, so an untrained cache predicts no radiance (and a frame then
  // shows the terminated paths' own terms, not noise from random weights).
  std::vector<float> w(kCacheParams, 0.0f);
  const int ins[kCacheLayers] = {48, 64, 64, 64, 64}, outs[kCacheLayers] = {64, 64, 64, 64, 16};
  uint32_t state = cfg.seed * 747796405u + 2891336453u;
  for (int l = 0; l < kCacheLayers; ++l) {
    const float bound = std::sqrt(6.0f / static_cast<float>(ins[l] + outs[l]));
    for (int o = 0; o < outs[l]; ++o)
      for (int i = 0; i < ins[l]; ++i) {
        state = state * 747796405u + 2891336453u;
        uint32_t h = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
        h = (h >> 22u) ^ h;
        const float u = static_cast<float>(h >> 8) * (1.0f / 16777216.0f);
        w[kCacheWeightOffset[l] + o * ins[l] + i] = l == 4 ? 0.0f : (2.0f * u - 1.0f) * bound;
      }
  }
  e = hipMemcpyAsync(c.w32, w.data(), sizeof(float) * kCacheParams, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(c.wm, 0, sizeof(float) * kCacheParams, s);
  if (e == hipSuccess) e = hipMemsetAsync(c.wv, 0, sizeof(float) * kCacheParams, s);
  if (e == hipSuccess) e = hipMemsetAsync(c.loss, 0, sizeof(float), s);
  if (e == hipSuccess) {
    k_cache_init_grid<<<stream_blocks(c.grid_entries), 256, 0, s>>>(c.grid32, c.grid16, c.gm, c.gv, c.ggrad,
                                                                   c.grid_entries, cfg.seed);
    k_cache_zero16<<<stream_blocks(c.grid_entries), 256, 0, s>>>(c.ggrad16, c.grid_entries);
    k_cache_pack<<<(kFwdHalves + kBwdHalves + 255) / 256, 256, 0, s>>>(c.w32, c.wf16, c.wb16);
    e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    err = std::string("radiance cache initialisation: ") + hipGetErrorString(e);
    cache_free(c);
    return false;
  }
  c.ready = true;
  return true;
}

void cache_infer(const CacheState& c, const float* queries, uint32_t n, float* out, hipStream_t s) {
  if (n == 0) return;
  const uint32_t per_block = 4 * 16 * kInferTiles;
  const uint32_t want = (n + per_block - 1) / per_block, cap = cu_count() * 4;
  k_cache_infer<<<want < cap ? want : cap, 256, 0, s>>>(grid_args(c), c.wf16, queries, n, nullptr, out, nullptr,
                                                        nullptr);
}

void cache_infer_add(const CacheState& c, const float4* queries, const float4* throughput, const uint32_t* count,
                     float4* radiance, hipStream_t s) {
  k_cache_infer<<<cu_count() * 4, 256, 0, s>>>(grid_args(c), c.wf16, reinterpret_cast<const float*>(queries), 0u,
                                               count, nullptr, throughput, radiance);
}

void cache_backprop(CacheState& c, const float* samples, uint32_t n, const uint32_t* n_dev, uint32_t first,
                    uint32_t stride, hipStream_t s) {
  const uint32_t cap = c.config.batch_capacity;
  const uint32_t rows = n_dev ? c.ld : ((n < cap ? n : cap) + 31u) & ~31u;
  if (rows == 0) return;
  const uint32_t want = (rows / 16 + 3) / 4, most = cu_count() * 2;
  k_cache_backprop<<<want < most ? want : most, 256, 0, s>>>(grid_args(c), c.wf16, c.wb16, samples, n, n_dev, first,
                                                             stride ? stride : 1u, cap, c.acts, c.deltas, c.ld,
                                                             c.loss);
  const uint32_t chunks = n_dev ? c.chunks : (rows + kChunk - 1) / kChunk;
  k_cache_wgrad<<<chunks, 256, 0, s>>>(c.acts, c.deltas, c.ld, n, n_dev, first, stride ? stride : 1u, cap,
                                        c.partials);
  k_cache_wsum<<<kCacheParams / 256, 256, 0, s>>>(c.partials, chunks, c.wgrad);
}

void cache_adam(CacheState& c, uint32_t n, const uint32_t* n_dev, uint32_t first, uint32_t stride, hipStream_t s) {
  ++c.step;
  const AdamArgs a = adam_args(c);
  const uint32_t cap = c.config.batch_capacity;
  stride = stride ? stride : 1u;
  k_cache_adam_w<<<kCacheParams / 256, 256, 0, s>>>(c.w32, c.wgrad, c.wm, c.wv, a, n, n_dev, first, stride, cap);
  k_cache_adam_grid<<<stream_blocks(c.grid_entries), 256, 0, s>>>(c.grid32, c.grid16, c.ggrad, c.ggrad16,
                                                                 c.dense_mask, c.config.log2_table + 1, c.gm, c.gv,
                                                                 c.grid_entries, a, n, n_dev, first, stride, cap);
  k_cache_pack<<<(kFwdHalves + kBwdHalves + 255) / 256, 256, 0, s>>>(c.w32, c.wf16, c.wb16);
}

void cache_sync_half(CacheState& c, hipStream_t s) {
  k_cache_pack<<<(kFwdHalves + kBwdHalves + 255) / 256, 256, 0, s>>>(c.w32, c.wf16, c.wb16);
  k_cache_half<<<stream_blocks(c.grid_entries), 256, 0, s>>>(c.grid32, c.grid16, c.grid_entries);
}

void cache_clear_grid_grad(CacheState& c, hipStream_t s) {
  k_cache_zero<<<stream_blocks(c.grid_entries), 256, 0, s>>>(c.ggrad, c.grid_entries);
  k_cache_zero16<<<stream_blocks(c.grid_entries), 256, 0, s>>>(c.ggrad16, c.grid_entries);
}

void cache_merge_grid_grad(CacheState& c, hipStream_t s) {
  k_cache_grad_merge<<<stream_blocks(c.grid_entries), 256, 0, s>>>(c.ggrad, c.ggrad16, c.dense_mask,
                                                                   c.config.log2_table + 1, c.grid_entries);
}

void cache_reset_moments(CacheState& c, hipStream_t s) {
  for (float* p : {c.gm, c.gv, c.ggrad})
    k_cache_zero<<<stream_blocks(c.grid_entries), 256, 0, s>>>(p, c.grid_entries);
  k_cache_zero16<<<stream_blocks(c.grid_entries), 256, 0, s>>>(c.ggrad16, c.grid_entries);
  for (float* p : {c.wm, c.wv}) k_cache_zero<<<kCacheParams / 256, 256, 0, s>>>(p, kCacheParams);
  c.step = 0;
}

}  // namespace lotus
