// Radiance cache: a multiresolution hash-grid encoding (FP16 features)
// feeding a 4 x 64 ReLU MLP, evaluated and trained on MFMA (FP16 operands,
// FP32 accumulation). No reference counterpart (a neural radiance cache is
// a roadmap candidate of the reference, docs/design/ROADMAP_POLICY.md:526-535);
// the network follows Mueller et al., "Instant Neural Graphics Primitives"
// (SIGGRAPH 2022) and "Real-time Neural Radiance Caching for Path Tracing"
// (SIGGRAPH 2021).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace lotus {

constexpr int kCacheLevels = 16;                                 // hash-grid levels
constexpr int kCacheFeatures = 2;                                // FP16 features per level
constexpr int kCacheGridInputs = kCacheLevels * kCacheFeatures;  // 32
constexpr int kCacheExtra = 16;   // direction, normal, roughness, base colour, metallic, 1, 4 zeros
constexpr int kCacheInputs = kCacheGridInputs + kCacheExtra;     // 48
constexpr int kCacheWidth = 64;   // hidden width
constexpr int kCacheLayers = 5;   // 4 hidden ReLU layers and a linear output
constexpr int kCacheOutputs = 16; // output rows (RGB used)
// Weight matrices [out][in], no biases: 64x48, 64x64 (three), 16x64.
constexpr int kCacheParams = 16384;
constexpr int kCacheWeightOffset[kCacheLayers + 1] = {0, 3072, 7168, 11264, 15360, 16384};
// A query: normalized position (3), unit direction towards the previous
// vertex (3), shading normal (3), roughness, base colour (3), metallic, 2
// unused: 16 floats. A training sample: a query, its RGB target, 1 unused.
constexpr int kQueryFloats = 16;
constexpr int kSampleFloats = 20;
// Per-sample output gradient clamp (keeps FP16 backward operands finite).
constexpr float kCacheDeltaClamp = 1024.0f;

struct CacheConfig {
  uint32_t log2_table = 19;
  uint32_t min_resolution = 16;
  uint32_t max_resolution = 2048;
  uint32_t batch_capacity = 65536;
  float learning_rate = 5e-3f;
  float beta1 = 0.9f, beta2 = 0.99f, epsilon = 1e-15f;
  uint32_t seed = 1;
};

struct CacheState {
  CacheConfig config;
  uint32_t resolution[kCacheLevels] = {};
  uint32_t dense_mask = 0;
  uint64_t grid_entries = 0;  // levels x table x features
  float* w32 = nullptr;       // master weights
  float* wgrad = nullptr;     // batch sums of the weight gradients
  float* wm = nullptr;
  float* wv = nullptr;
  _Float16* wf16 = nullptr;   // forward operand image (kCacheParams halves)
  _Float16* wb16 = nullptr;   // backward operand image
  float* grid32 = nullptr;    // [level][entry][feature] master
  _Float16* grid16 = nullptr;
  float* ggrad = nullptr;      // gradient sums of the dense levels (FP32)
  _Float16* ggrad16 = nullptr; // gradient sums of the hashed levels (FP16 pairs)
  float* gm = nullptr;
  float* gv = nullptr;
  _Float16* acts = nullptr;   // layer inputs, [row][sample], rows 48 + 4 x 64
  _Float16* deltas = nullptr; // layer output gradients, rows 4 x 64 + 16
  float* partials = nullptr;  // weight-gradient partial sums per sample chunk
  float* loss = nullptr;      // batch sum of the loss
  uint32_t* count = nullptr;  // device sample count for host-sized launches
  uint32_t ld = 0;            // row length of acts / deltas (capacity, padded)
  uint32_t chunks = 0;
  uint32_t step = 0;
  bool ready = false;
};

bool cache_init(CacheState& c, const CacheConfig& cfg, hipStream_t s, std::string& err);
void cache_free(CacheState& c);
// out[n][3] = network(queries[n][16]) (device pointers).
void cache_infer(const CacheState& c, const float* queries, uint32_t n, float* out, hipStream_t s);
// For the first *count queries: radiance[slot].xyz += throughput.xyz * max(prediction, 0),
// slot = throughput.w's bits.
void cache_infer_add(const CacheState& c, const float4* queries, const float4* throughput, const uint32_t* count,
                     float4* radiance, hipStream_t s);
// Gradient sums of the relative L2 loss over samples first + stride * i of
// the first `total` samples, total = n, or min(*n_dev, n) when n_dev is given
// (n then being the sample buffer's capacity), at most the batch capacity of
// them, into wgrad / ggrad, and the loss sum into c.loss.
void cache_backprop(CacheState& c, const float* samples, uint32_t n, const uint32_t* n_dev, uint32_t first,
                    uint32_t stride, hipStream_t s);
// One Adam step on the gradient sums divided by the sample count; clears the
// grid gradients. Skipped on the device when the count is zero.
void cache_adam(CacheState& c, uint32_t n, const uint32_t* n_dev, uint32_t first, uint32_t stride, hipStream_t s);
// Rewrites the FP16 operand images and grid copy from the FP32 masters.
void cache_sync_half(CacheState& c, hipStream_t s);
void cache_clear_grid_grad(CacheState& c, hipStream_t s);
// The hashed levels' FP16 gradient sums into ggrad, which then holds every level.
void cache_merge_grid_grad(CacheState& c, hipStream_t s);
// Zeroes the Adam moments and grid gradients and restarts the step count.
void cache_reset_moments(CacheState& c, hipStream_t s);

}  // namespace lotus
