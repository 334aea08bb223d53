// Device-side vector arithmetic, random numbers and ray helpers for the gfx950
// wavefront path tracer. Every float expression is evaluated in IEEE single
// precision in the order written: the library is compiled with
// -ffp-contract=off, and `/` and sqrtf are correctly rounded (HIP's default
// -fhip-fp32-correctly-rounded-divide-sqrt), so decisions made from these
// values are reproducible on any IEEE machine.
//
// Replacements for functions whose rounding the reference leaves to its GPU:
//   sin/cos(2 pi u) -> SinCosTurns (quadrant reduction, Taylor to degree 9/10);
//   rsqrt(x) -> 1 / sqrt(x);  normalize(v) -> v * rsqrt(dot(v, v));
//   pow(x, 5) -> ((x*x)*(x*x))*x;  dot -> (x*x' + y*y') + z*z'.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lotus {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kFloatMax = 3.402823466e+38f;
constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr uint32_t kNoLights = 0xFFFFFFFFu;

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float rsqrt_rn(float x) { return 1.0f / sqrtf(x); }
__device__ __forceinline__ float length3(V3 v) { return sqrtf(dot3(v, v)); }
__device__ __forceinline__ V3 normalize3(V3 v) { return v * rsqrt_rn(dot3(v, v)); }
// max/min with std::max/std::min semantics: the first operand unless the
// second compares greater (smaller).
__device__ __forceinline__ float maxf(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float minf(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float max3(V3 v) { return maxf(v.x, maxf(v.y, v.z)); }
__device__ __forceinline__ bool finite3(V3 v) {
  return isfinite(v.x) && isfinite(v.y) && isfinite(v.z);
}
__device__ __forceinline__ float saturate(float x) { return minf(maxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float comp(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ __forceinline__ V3 xyz(float4 f) { return V3{f.x, f.y, f.z}; }

// Column-major 4x4 times (x, y, z, w), row r: ((m[r]x + m[4+r]y) + m[8+r]z) + m[12+r]w.
__device__ __forceinline__ void mul_point(const float* m, float x, float y, float z, float w,
                                          float* out) {
#pragma unroll
  for (int r = 0; r < 4; ++r) out[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
}

// sin and cos of 2 pi u, u in [0, 1].
__device__ __forceinline__ void sincos_turns(float u, float* s, float* c) {
  const float t = u * 4.0f;
  int q = static_cast<int>(t);
  float f = t - static_cast<float>(q);
  if (f > 0.5f) {
    f = f - 1.0f;
    q = q + 1;
  }
  const float a = f * 1.57079632679489662f;
  const float a2 = a * a;
  const float sn = a + a * a2 * (-1.66666666666666667e-1f +
      a2 * (8.33333333333333333e-3f + a2 * (-1.98412698412698413e-4f +
      a2 * 2.75573192239858907e-6f)));
  const float cs = 1.0f + a2 * (-0.5f + a2 * (4.16666666666666667e-2f +
      a2 * (-1.38888888888888889e-3f + a2 * (2.48015873015873016e-5f +
      a2 * -2.75573192239858907e-7f))));
  const int k = q & 3;
  *s = k == 0 ? sn : (k == 1 ? cs : (k == 2 ? -sn : -cs));
  *c = k == 0 ? cs : (k == 1 ? -sn : (k == 2 ? -cs : sn));
}

// atan2(y, x) / pi in [-1, 1]: reduction to |t| <= tan(pi/8) (swap of the
// arguments, then atan(r) = pi/4 + atan((r - 1) / (r + 1))) and the Taylor
// series of atan to t^15, error below 2e-8 rad. It replaces the reference
// GPU's atan2 and acos (scene.slang:805-809) so the oracle matches bit for bit.
__device__ __forceinline__ float atan2_pi(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  if (ax == 0.0f && ay == 0.0f) return 0.0f;
  const bool swap = ay > ax;
  const float r = swap ? ax / ay : ay / ax;
  const bool shift = r > 0.41421356237309505f;
  const float t = shift ? (r - 1.0f) / (r + 1.0f) : r;
  const float t2 = t * t;
  const float p = t + t * t2 * (-3.33333333333333333e-1f + t2 * (2.0e-1f + t2 * (-1.42857142857142857e-1f +
      t2 * (1.11111111111111111e-1f + t2 * (-9.09090909090909091e-2f + t2 * (7.69230769230769231e-2f +
      t2 * -6.66666666666666667e-2f))))));
  float a = (shift ? 0.785398163397448310f + p : p) * 0.318309886183790672f;
  if (swap) a = 0.5f - a;
  if (x < 0.0f) a = 1.0f - a;
  return y < 0.0f ? -a : a;
}
// acos(y) / pi for y in [-1, 1].
__device__ __forceinline__ float acos_pi(float y) {
  return atan2_pi(sqrtf(fmaxf(0.0f, (1.0f - y) * (1.0f + y))), y);
}

// PCG RXS-M-XS hash and generator (sampling.slang:9-25).
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
  const uint32_t state = v * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
__device__ __forceinline__ uint32_t seed_random(uint32_t pixel, uint32_t sample) {
  return pcg_hash(pixel ^ pcg_hash(sample));
}
__device__ __forceinline__ float next_random(uint32_t& state) {
  state = state * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return static_cast<float>(((word >> 22u) ^ word) >> 8u) * (1.0f / 16777216.0f);
}

// scene.slang:1454-1469 (Waechter and Binder, Ray Tracing Gems 2019).
__device__ __forceinline__ float offset_axis(float p, float n) {
  const int offset = static_cast<int>(256.0f * n);
  const float moved = __int_as_float(static_cast<int>(
      static_cast<uint32_t>(__float_as_int(p)) + static_cast<uint32_t>(p < 0.0f ? -offset : offset)));
  return fabsf(p) < (1.0f / 32.0f) ? p + (1.0f / 65536.0f) * n : moved;
}
__device__ __forceinline__ V3 offset_ray_origin(V3 p, V3 n) {
  return V3{offset_axis(p.x, n.x), offset_axis(p.y, n.y), offset_axis(p.z, n.z)};
}

// Wave-aggregated append: one atomic per wave, lanes ordered by lane id.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool active) {
  const uint64_t mask = __ballot(active);
  const uint32_t lane = __lane_id();
  uint32_t base = 0;
  if (mask == 0) return 0;
  const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(mask)) - 1);
  if (lane == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
  base = __shfl(base, static_cast<int>(leader));
  return base + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
}

// Wave-aggregated counter increment.
__device__ __forceinline__ void wave_count(uint32_t* counter, bool active) {
  const uint64_t mask = __ballot(active);
  if (mask != 0 && __lane_id() == static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(mask)) - 1))
    atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
}

// Increments of counters that may differ between lanes: one atomic per
// distinct counter among the active lanes (a few per wave where lanes are at
// different bounces), instead of one per lane on a handful of addresses.
__device__ __forceinline__ void wave_count_each(uint32_t* counter, bool active) {
  uint64_t todo = __ballot(active);
  const uint32_t lane = __lane_id();
  const unsigned long long mine = reinterpret_cast<unsigned long long>(counter);
  while (todo != 0) {
    const int leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
    const unsigned long long theirs = __shfl(mine, leader);
    const uint64_t same = __ballot(active && mine == theirs) & todo;
    if (lane == static_cast<uint32_t>(leader)) atomicAdd(counter, static_cast<uint32_t>(__popcll(same)));
    todo &= ~same;
  }
}

// Each lane of a fully active wave reserves `k` consecutive slots of
// *counter and gets its first one: a wave-wide prefix sum and one atomic.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counter, uint32_t k) {
  const uint32_t lane = __lane_id();
  uint32_t incl = k;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t v = __shfl_up(incl, d);
    if (lane >= static_cast<uint32_t>(d)) incl += v;
  }
  const uint32_t total = __shfl(incl, 63);
  uint32_t base = 0;
  if (lane == 63 && total != 0) base = atomicAdd(counter, total);
  base = __shfl(base, 63);
  return base + incl - k;
}

}  // namespace lotus
