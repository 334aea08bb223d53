// The wavefront integrator's kernels for gfx950 (the reference's Camera /
// Intersect / Shade / Tail kernels, backend/vulkan/shaders/wavefront/
// wavefront.slang:206-309, and its resolve pass, path_trace.slang:106-139),
// plus the pixel-centre camera pass (path_trace.slang:69-97) and a plain
// trace kernel over caller rays.
//
// Scheduling differs from the reference by design. HIP has no indirect
// dispatch, so every kernel is a persistent grid, sized from its occupancy to
// fill every CU, striding over a queue whose length it reads from device
// memory. Queues are structure-of-arrays records in queue order, appended with
// one atomic per wave (ballot + popcount prefix), so a kernel's loads of its
// queue are contiguous. One traversal launch per bounce takes both the
// extension rays of bounce b and the shadow rays that shading at bounce b-1
// sent: a shadow ray that reaches its light leaves its contribution pending in
// its slot, and an escaping ray leaves its environment term pending; shading
// at the next bounce, the tail or the resolve adds them, so every path's
// radiance is summed in the reference's order (emission, direct light, next
// emission ..., environment).
#include "temporal.hpp"

namespace lotus {
namespace {

// Adds per-lane traversal counts to stats[2 kind], stats[2 kind + 1] (wave
// sum, one atomic per wave).
__device__ __forceinline__ void flush_stats(unsigned long long* stats, int kind, uint32_t v, uint32_t t) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v += __shfl_xor(v, off);
    t += __shfl_xor(t, off);
  }
  if (__lane_id() == 0 && (v | t) != 0) {
    atomicAdd(stats + 2 * kind, static_cast<unsigned long long>(v));
    atomicAdd(stats + 2 * kind + 1, static_cast<unsigned long long>(t));
  }
}

// Traversal counts of one kind kept per lane in the LDS word just past the
// lane's stack ring (its workgroup's ring is declared one entry deeper): no
// atomics, and no register beyond the stack's own address. Visits count in
// the low half, tests in the high half. Each wave adds its sums to
// stats[2 kind], stats[2 kind + 1] at its end.
struct LaneStats {
  TraceStack::LdsEntry* word;

  __device__ __forceinline__ void init(TraceStack& st) {
    word = st.mine + kRing * kTraceBlock;
    *word = 0;
  }
  __device__ __forceinline__ void add(uint32_t v, uint32_t t) {
    *word += static_cast<unsigned long long>(v) | (static_cast<unsigned long long>(t) << 32);
  }
  __device__ __forceinline__ void flush(unsigned long long* stats, int kind) {
    const unsigned long long w = *word;
    flush_stats(stats, kind, static_cast<uint32_t>(w), static_cast<uint32_t>(w >> 32));
  }
};

// Traversal counts of a workgroup, kept in LDS rather than in the walkers'
// registers (the four per-lane sums of the combined traversal kernel cost it
// a wave of occupancy per SIMD): a lane adds its finished walk's counts with
// LDS atomics, and the workgroup adds its totals to stats[2 (first + k)],
// stats[2 (first + k) + 1] once, at its end. Every thread of the workgroup
// must call init() and flush().
template <int kKinds>
struct BlockStats {
  unsigned long long* c;  // [2 kKinds] in LDS

  __device__ __forceinline__ void init(unsigned long long* lds) {
    c = lds;
    if (threadIdx.x < 2 * kKinds) c[threadIdx.x] = 0;
    __syncthreads();
  }
  __device__ __forceinline__ void add(int kind, uint32_t v, uint32_t t) {
    atomicAdd(c + 2 * kind, static_cast<unsigned long long>(v));
    atomicAdd(c + 2 * kind + 1, static_cast<unsigned long long>(t));
  }
  __device__ __forceinline__ void flush(unsigned long long* stats, int first) {
    __syncthreads();
    if (threadIdx.x < 2 * kKinds && c[threadIdx.x] != 0) atomicAdd(stats + 2 * first + threadIdx.x, c[threadIdx.x]);
  }
};

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t gstride() { return gridDim.x * blockDim.x; }

// Walks with hand-over from fixed sequences: lane g takes items g, g + G,
// g + 2G, ... (G lanes in the grid) like the strided walks, but a lane whose
// walk ends does not wait for the slowest ray of its wave: once a quarter of
// the wave (or all of it) is ready, the ready lanes hand their results to
// finish(item, walk) together and start their next items. No atomics: the
// sequences are fixed, so neighbouring lanes keep neighbouring items.
template <class Cov, class Fetch, class Finish>
__device__ __forceinline__ void sequence_walks(const SceneView& S, TraceStack& st, uint32_t total,
                                               Fetch fetch, Finish finish, const Cov& cov) {
  WalkT<Cov> w;
  w.cov = cov;
  bool working = false;  // a walk in progress
  bool holding = false;  // a finished walk whose result awaits hand-over
  uint32_t item = 0;
  uint32_t next = gtid();
  const uint32_t stride = gstride();
  while (true) {
    const uint64_t active = __ballot(true);
    const uint64_t ready = __ballot(!working);
    if (__ballot(working || holding || next < total) == 0) break;
    if (ready != 0 && (__popcll(ready) >= kRefill || ready == active)) {
      if (holding) {
        finish(item, w);
        holding = false;
      }
      if (!working && next < total) {
        item = next;
        next += stride;
        fetch(item, w);
        st.clear();
        // an item with nothing to walk (fetch may start a finished walk)
        // goes straight to hand-over: a step of a finished walk would read
        // its end marker as a leaf
        if (w.done()) holding = true;
        else working = true;
      }
    }
    if (working) {
      w.step(S.nodes, S.tris, st);
      if (w.done()) {
        working = false;
        holding = true;
      }
    }
  }
}

// The same walks without hand-over: each lane follows one ray to its end,
// striding over the queue.
template <class Cov, class Fetch, class Finish>
__device__ __forceinline__ void strided_walks(const SceneView& S, TraceStack& st, uint32_t total, Fetch fetch,
                                              Finish finish, const Cov& cov) {
  for (uint32_t i = gtid(); i < total; i += gstride()) {
    WalkT<Cov> w;
    w.cov = cov;
    fetch(i, w);
    st.clear();
    w.run(S.nodes, S.tris, st);
    finish(i, w);
  }
}

template <bool kDynamic, class Cov, class Fetch, class Finish>
__device__ __forceinline__ void walks(const SceneView& S, TraceStack& st, uint32_t total, Fetch fetch,
                                      Finish finish, const Cov& cov) {
  if constexpr (kDynamic) sequence_walks(S, st, total, fetch, finish, cov);
  else strided_walks(S, st, total, fetch, finish, cov);
}

// The coverage policy of a kernel variant: the scene's when kCov, else none.
template <bool kCov>
using CoverageOf = typename std::conditional<kCov, SceneCoverage, NoCoverage>::type;
template <bool kCov>
__device__ __forceinline__ CoverageOf<kCov> coverage_of(const SceneView& S) {
  if constexpr (kCov) return SceneCoverage{&S};
  else return NoCoverage{};
}

// A shadow ray's pending light once its walk ended (ShadowedLight's end,
// transport.slang:105-114): without coverage, the contribution shading
// stored, when nothing blocked the ray; with coverage, f * radiance *
// (transmittance * weight / pdf) from the factors shading stored, times the
// path's throughput, when the transmittance is positive.
template <class Cov>
__device__ __forceinline__ bool shadow_light(const Queues& q, uint32_t j, const WalkT<Cov>& w, float4& nee) {
  if constexpr (!Cov::kEnabled) {
    if (w.hit()) return false;
    const float4 c4 = q.sh_c[j];
    nee = make_float4(c4.x, c4.y, c4.z, 1.0f);
    return true;
  } else {
    const float T = w.transmittance();
    if (!(T > 0.0f)) return false;
    const float4 f4 = q.sh_f[j], t4 = q.sh_t[j];
    const V3 d = shadowed_delta(V3{f4.x, f4.y, f4.z}, f4.w, t4.w, V3{t4.x, t4.y, t4.z}, T);
    nee = make_float4(d.x, d.y, d.z, 1.0f);
    return true;
  }
}

__device__ __forceinline__ uint32_t* ray_count(uint32_t* c, uint32_t b) { return c + counter_word(b, 0); }
__device__ __forceinline__ uint32_t* hit_count(uint32_t* c, uint32_t b) { return c + counter_word(b, 1); }
__device__ __forceinline__ uint32_t* shadow_count(uint32_t* c, uint32_t b) { return c + counter_word(b, 2); }

__device__ __forceinline__ void push_hit(const Queues& q, uint32_t k, uint32_t slot, V3 o, V3 d,
                                         const HitRecord& h) {
  q.hit_o[k] = make_float4(o.x, o.y, o.z, __uint_as_float(slot));
  q.hit_d[k] = make_float4(d.x, d.y, d.z, __uint_as_float(h.tri));
  q.hit_h[k] = make_uint4(h.instance, h.primitive, __float_as_uint(h.u), __float_as_uint(h.v));
}

// wavefront.slang:206-225: a camera path per pixel of the data window, its
// camera ray traced at once; a hit is queued for shading, a miss leaves no
// sample. Clears the slot's pending terms.
template <bool kDynamic, bool kCov>  // kCov: the scene has covered surfaces
__global__ void __launch_bounds__(kTraceBlock) k_camera(SceneView S, FrameConst F, Queues q, uint32_t* counters,
                                                        uint2* spill, int spill_cap, unsigned long long* stats,
                                                        uint32_t tile_shift) {
  __shared__ uint2 ring[(kRing + 1) * kTraceBlock];  // + the lanes' traversal counts
  TraceStack st;
  st.init(ring, spill, spill_cap);
  const uint32_t rw = F.rect[2], rh = F.rect[3];
  if (gtid() == 0) *ray_count(counters, 0) = rw * rh;
  // Items walk the data window in pixel tiles of 64, a wave's rays one tile
  // (2^tile_shift wide: 8 x 8, 16 x 4 or 4 x 16), so a wave's rays stay
  // close together through the tree (and so do the hits and the next rays
  // they queue). Items past the window's edge in its last tiles are skipped.
  const uint32_t tw = 1u << tile_shift, th = 64u >> tile_shift;
  const uint32_t tiles_x = (rw + tw - 1) / tw, tiles = tiles_x * ((rh + th - 1) / th);
  const auto pixel = [&](uint32_t i, uint32_t& x, uint32_t& y) {
    const uint32_t tile = i >> 6, within = i & 63u;
    x = (tile % tiles_x) * tw + (within & (tw - 1u));
    y = (tile / tiles_x) * th + (within >> tile_shift);
    return x < rw && y < rh;
  };
  LaneStats ls;
  ls.init(st);
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  walks<kDynamic>(
      S, st, tiles * 64,
      [&](uint32_t i, auto& w) {
        uint32_t rng, x, y;
        if (!pixel(i, x, y)) {
          w.start(Ray{V3{0.0f, 0.0f, 0.0f}, 0.0f, V3{0.0f, 0.0f, 1.0f}, 0.0f}, false);
          w.cur = kWalkDone;
          return;
        }
        w.start(start_path(F, F.rect[0] + x, F.rect[1] + y, frame_sample_index(F), rng), false);
      },
      [&](uint32_t i, const auto& w) {
        uint32_t x, y;
        if (!pixel(i, x, y)) return;
        x += F.rect[0];
        y += F.rect[1];
        const uint32_t slot = y * F.width + x;
        const bool hit = w.hit();
        ls.add(w.visits, w.tests);
        const uint32_t k = wave_append(hit_count(counters, 0), hit);
        if (hit) {
          uint32_t rng;
          const Ray r = start_path(F, x, y, frame_sample_index(F), rng);  // again: the ray, and the path's random state
          push_hit(q, k, slot, r.o, r.d, w.record(S.tris));
          q.rad[slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(rng));
        } else {
          q.rad[slot] = zero;
        }
        q.nee[slot] = zero;
        q.esc[slot] = zero;
        if (F.light_sampling == kLightRestir) {
          // restir.slang:171-176: a failure of an earlier sample of the
          // accumulation stays flagged
          const uint32_t failed = F.accumulated != 0 ? (__float_as_uint(q.restir[4ull * slot + 1].w) & 2u) : 0u;
          q.restir[4ull * slot] = make_float4(__uint_as_float(kNoReservoirSample), 0.0f, 0.0f, 0.0f);
          q.restir[4ull * slot + 1] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(failed));
          q.restir[4ull * slot + 2] = zero;
        }
      },
      coverage_of<kCov>(S));
  ls.flush(stats, 0);
}

// restir.slang:163-233 (WriteInitialReservoir) over the first hits: each
// draws restir_candidates lights from its own random stream (seeded from the
// pixel and sample ^ 0xD1B54A35, so the path's own draws are unchanged),
// streams them by the maximum component of their unshadowed contribution,
// and traces one visibility ray for the selection (counted apart from the
// queued shadow rays).
template <bool kLights, bool kGltf = false>  // extras (lights, images, lookups); glTF PBR surfaces
__global__ void __launch_bounds__(kTraceBlock) k_restir_initial(SceneView S, FrameConst F, Queues q,
                                                                uint32_t* counters, uint2* spill, int spill_cap,
                                                                unsigned long long* stats) {
  __shared__ uint2 ring[(kRing + 1) * kTraceBlock];  // + the lanes' traversal counts
  TraceStack st;
  st.init(ring, spill, spill_cap);
  LaneStats ls_counts;  // in LDS: the kernel sits one register above an occupancy step
  ls_counts.init(st);
  const uint32_t count = *hit_count(counters, 0);
  for (uint32_t i = gtid(); i < count; i += gstride()) {
    const float4 o4 = q.hit_o[i], d4 = q.hit_d[i];
    const uint4 h4 = q.hit_h[i];
    const uint32_t slot = __float_as_uint(o4.w);
    const V3 dir = xyz(d4);
    const Surface sf = reconstruct_hit<kLights, kGltf>(S, __float_as_uint(d4.w), h4.x, __uint_as_float(h4.z),
                                                __uint_as_float(h4.w), dir);
    const V3 n = sf.shading_normal;
    const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    const V3 ft = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    const V3 fb = V3{b, sign + n.y * n.y * a, -n.y};
    const V3 md = -dir;
    const V3 wo = V3{dot3(md, ft), dot3(md, fb), dot3(md, n)};
    const V3 origin = offset_ray_origin(sf.position, sf.normal);
    Reservoir r = empty_reservoir();
    float ru = 0.0f, rv = 0.0f, rproposal = 0.0f;
    uint32_t flags = 1u | (F.accumulated != 0 ? (__float_as_uint(q.restir[4ull * slot + 1].w) & 2u) : 0u);
    uint32_t rng = seed_random(slot, F.sample_index ^ 0xD1B54A35u);
    if (F.max_bounces != 0 && S.light_count != 0) {
      for (uint32_t k = 0; k < F.restir_candidates; ++k) {
        const float cx = next_random(rng), cy = next_random(rng);
        const float u = next_random(rng), v = next_random(rng);
        const float random = next_random(rng);
        const uint32_t light = choose_reservoir_light(S, F.restir_proposal, cx, cy);
        const float proposal = reservoir_proposal_pdf(S, F.restir_proposal, light);
        const LightSample ls = reconstruct_reservoir_light<kLights>(S, light, u, v, origin);
        const V3 f = candidate_contribution(sf, ft, fb, wo, ls, S.lights[light].pdf);
        bool selected = false;
        if (!isfinite(ls.pdf) || !finite3(f) || !stream_candidate(r, light, max3(f), proposal, random, selected)) {
          flags |= 2u;
          break;
        }
        if (selected) {
          ru = u;
          rv = v;
          rproposal = proposal;
        }
      }
    }
    float4 contribution = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const bool trace = (flags & 2u) == 0u && r.weight_sum > 0.0f;
    wave_count(restir_ray_count(counters), trace);
    if (trace) {
      const LightSample ls = reconstruct_reservoir_light<kLights>(S, r.sample, ru, rv, origin);
      uint32_t visits = 0, tests = 0;
      const float visibility = walk_transmittance(S.nodes, S.tris, st, Ray{origin, 0.0f, ls.direction, ls.distance},
                                                  visits, tests, coverage_of<kLights>(S));
      ls_counts.add(visits, tests);
      const float normalization = reservoir_normalization(r);
      const V3 estimate = candidate_contribution(sf, ft, fb, wo, ls, S.lights[r.sample].pdf) *
                          (visibility * normalization);
      if (!isfinite(normalization) || !finite3(estimate))
        flags |= 2u;
      else
        contribution = make_float4(estimate.x, estimate.y, estimate.z, visibility);
    }
    q.restir[4ull * slot] = make_float4(__uint_as_float(r.sample), r.weight_sum, r.selected_target,
                                        __uint_as_float(r.count));
    q.restir[4ull * slot + 1] = make_float4(ru, rv, rproposal, __uint_as_float(flags));
    q.restir[4ull * slot + 2] = contribution;
    if (F.restir_surf) {
      float view[4];
      mul_point(F.view, sf.position.x, sf.position.y, sf.position.z, 1.0f, view);
      store_reuse_surface(F.restir_surf + kReuseSurfaceWords * slot, sf, origin, -view[2] / view[3], wo,
                          stable_instance_id(F, h4.x), stable_material_id(F, S.instance_material, h4.x));
    }
  }
  ls_counts.flush(stats, 2);
}

// Every pixel's initial record emptied (temporal frames: the reuse passes
// read records outside the data window too), and with `surf` every
// receiver surface zeroed, as restir.slang:306-307 starts each receiver
// empty (a pixel whose camera ray misses keeps the zero surface).
__global__ void __launch_bounds__(256) k_restir_clear(float4* restir, float4* surf, uint32_t pixels) {
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (uint32_t i = gtid(); i < pixels; i += gstride()) {
    restir[4ull * i] = make_float4(__uint_as_float(kNoReservoirSample), 0.0f, 0.0f, 0.0f);
    restir[4ull * i + 1] = zero;
    restir[4ull * i + 2] = zero;
    restir[4ull * i + 3] = zero;
    if (surf)
      for (uint32_t k = 0; k < kReuseSurfaceWords; ++k) surf[kReuseSurfaceWords * i + k] = zero;
  }
}

// A reuse record in registers: the reservoir record and {normalization,
// supporting candidates, previous index, age}.
struct ReuseRecord {
  Reservoir r;
  float u, v, proposal;
  uint32_t flags;
  float4 contribution;
  float normalization;
  uint32_t supporting, previous, age;
};

__device__ __forceinline__ ReuseRecord load_reuse(const float4* p) {
  const float4 a = p[0], b = p[1], c = p[2], d = p[3];
  ReuseRecord x;
  x.r = Reservoir{__float_as_uint(a.x), a.y, a.z, __float_as_uint(a.w)};
  x.u = b.x;
  x.v = b.y;
  x.proposal = b.z;
  x.flags = __float_as_uint(b.w);
  x.contribution = c;
  x.normalization = d.x;
  x.supporting = __float_as_uint(d.y);
  x.previous = __float_as_uint(d.z);
  x.age = __float_as_uint(d.w);
  return x;
}

__device__ __forceinline__ void store_reuse(float4* p, const ReuseRecord& x) {
  p[0] = make_float4(__uint_as_float(x.r.sample), x.r.weight_sum, x.r.selected_target, __uint_as_float(x.r.count));
  p[1] = make_float4(x.u, x.v, x.proposal, __uint_as_float(x.flags));
  p[2] = x.contribution;
  p[3] = make_float4(x.normalization, __uint_as_float(x.supporting), __uint_as_float(x.previous),
                     __uint_as_float(x.age));
}

// reservoir.slang:64-97 (MergeReevaluatedReservoir).
__device__ __forceinline__ bool merge_reevaluated(Reservoir& d, const Reservoir& src, float normalization,
                                                  float target, uint32_t cap, float random, bool& selected) {
  selected = false;
  const uint32_t count = src.count < cap ? src.count : cap;
  const bool zero = src.weight_sum == 0.0f && src.sample == kNoReservoirSample && src.selected_target == 0.0f &&
                    normalization == 0.0f && target == 0.0f;
  const bool positive = src.weight_sum > 0.0f && src.selected_target > 0.0f && src.sample != kNoReservoirSample &&
                        src.count != 0u && normalization > 0.0f;
  const bool valid_destination = d.weight_sum == 0.0f
                                     ? d.sample == kNoReservoirSample && d.selected_target == 0.0f
                                     : d.weight_sum > 0.0f && d.selected_target > 0.0f &&
                                           d.sample != kNoReservoirSample && d.count != 0u;
  if (!valid_destination || !isfinite(d.weight_sum) || !isfinite(d.selected_target) || !isfinite(src.weight_sum) ||
      !isfinite(src.selected_target) || !(zero || positive) || !isfinite(normalization) || !isfinite(target) ||
      target < 0.0f || cap == 0u || !isfinite(random) || random < 0.0f || random >= 1.0f ||
      count > 0xFFFFFFFFu - d.count)
    return false;
  const float weight = (target * normalization) * static_cast<float>(count);
  const float sum = d.weight_sum + weight;
  if (!isfinite(weight) || !isfinite(sum) || (target > 0.0f && weight == 0.0f)) return false;
  selected = weight > 0.0f && (d.weight_sum == 0.0f || random < weight / sum);
  d.weight_sum = sum;
  d.count += count;
  if (selected) {
    d.sample = src.sample;
    d.selected_target = target;
  }
  return true;
}

// reservoir.slang:99-103.
__device__ __forceinline__ float reuse_normalization(const Reservoir& r, uint32_t supporting) {
  return r.weight_sum > 0.0f ? (r.weight_sum / static_cast<float>(supporting)) / r.selected_target : 0.0f;
}

__device__ __forceinline__ bool same_sample(const ReuseRecord& a, const ReuseRecord& b) {
  return a.r.sample == b.r.sample && a.u == b.u && a.v == b.v;
}

template <bool kLights>
__device__ __forceinline__ V3 reuse_contribution_uv(const SceneView& S, const ReuseSurface& r, uint32_t light, float u,
                                                    float v) {
  return reuse_contribution(S, r, light, reconstruct_reservoir_light<kLights>(S, light, u, v, r.origin));
}

// restir.slang:278-287 (HistoryTarget): the selection's target at a source's
// surface, stored when the source selected the same sample.
template <bool kLights>
__device__ __forceinline__ float history_target(const SceneView& S, const ReuseRecord& source,
                                                const ReuseSurface& surface, const ReuseRecord& selected,
                                                bool& valid) {
  if (same_sample(source, selected)) {
    valid = isfinite(source.r.selected_target);
    return source.r.selected_target;
  }
  const V3 c = reuse_contribution_uv<kLights>(S, surface, selected.r.sample, selected.u, selected.v);
  valid = finite3(c);
  return max3(c);
}

// The final estimate at the receiver (restir.slang:289-300, 372-385):
// visibility reused from the same pixel's record when it selected the same
// sample, else traced.
template <bool kLights>
__device__ __forceinline__ void resolve_estimate(const SceneView& S, TraceStack& st, const ReuseSurface& receiver,
                                                 ReuseRecord& result, const ReuseRecord& own, uint32_t& rays,
                                                 uint32_t& visits, uint32_t& tests) {
  if ((result.flags & 2u) != 0u || !(result.r.weight_sum > 0.0f)) return;
  const LightSample ls = reconstruct_reservoir_light<kLights>(S, result.r.sample, result.u, result.v, receiver.origin);
  float visibility;
  if (same_sample(result, own)) {
    visibility = own.contribution.w;
  } else {
    ++rays;
    visibility = walk_transmittance(S.nodes, S.tris, st, Ray{receiver.origin, 0.0f, ls.direction, ls.distance},
                                    visits, tests, coverage_of<kLights>(S));
  }
  const V3 estimate = reuse_contribution(S, receiver, result.r.sample, ls) * (visibility * result.normalization);
  if (!isfinite(result.normalization) || result.normalization <= 0.0f || !finite3(estimate))
    result.flags |= 2u;
  else
    result.contribution = make_float4(estimate.x, estimate.y, estimate.z, visibility);
}

// restir.slang:302-387 (WriteTemporalReservoir) per pixel of the target.
template <bool kLights, bool kGltf = false>  // analytic lights or an environment image; glTF PBR receivers
__global__ void __launch_bounds__(kTraceBlock) k_restir_temporal(SceneView S, FrameConst F, RestirReuseArgs A,
                                                                 uint32_t* counters, uint2* spill, int spill_cap,
                                                                 unsigned long long* stats) {
  __shared__ uint2 ring[kRing * kTraceBlock];
  TraceStack st;
  st.init(ring, spill, spill_cap);
  const uint32_t pixels = F.width * F.height;
  uint32_t visits = 0, tests = 0;
  for (uint32_t slot = gtid(); slot < pixels; slot += gstride()) {
    const ReuseRecord init = load_reuse(A.initial + 4ull * slot);
    ReuseRecord result = init;
    result.previous = 0xFFFFFFFFu;
    result.age = 0u;
    result.normalization = reservoir_normalization(init.r);
    result.supporting = init.r.count;
    uint32_t rays = 0;
    if ((init.flags & 1u) != 0u) {
      const ReuseSurface receiver = load_reuse_surface<kGltf>(A.surf + kReuseSurfaceWords * slot);
      // TemporalSampleHistory (temporal.slang:73-79): a valid reprojection
      // onto the same placement as this receiver
      const bool history = A.history != 0u && (result.flags & 2u) == 0u && A.validity[slot].valid != 0u &&
                           A.tsurf[slot].instance_id == receiver.instance_id &&
                           A.tsurf[slot].material_id == receiver.material_id;
      if (history) {
        const TemporalValidityPixel vd = A.validity[slot];
        const uint32_t pi = vd.previous_index;
        const ReuseRecord src = load_reuse(A.prev_hist + 4ull * pi);
        const ReuseSurface psurf = load_reuse_surface<kGltf>(A.prev_surf + kReuseSurfaceWords * pi);
        if ((src.flags & 1u) != 0u && (src.flags & 2u) == 0u && psurf.instance_id == receiver.instance_id &&
            psurf.material_id == receiver.material_id) {
          const uint32_t count = src.r.count < A.max_history ? src.r.count : A.max_history;
          V3 received = V3{0.0f, 0.0f, 0.0f};
          if (src.r.weight_sum > 0.0f) {
            if (src.r.sample >= S.light_count) result.flags |= 2u;
            else received = reuse_contribution_uv<kLights>(S, receiver, src.r.sample, src.u, src.v);
          }
          uint32_t rng = seed_random(slot, F.sample_index ^ 0x94D049BBu);
          bool selected = false;
          if (!finite3(received) || (result.flags & 2u) != 0u ||
              !merge_reevaluated(result.r, src.r, src.normalization, max3(received), A.max_history,
                                 next_random(rng), selected)) {
            result.flags |= 2u;
          } else {
            result.previous = pi;
            if (selected) {
              result.u = src.u;
              result.v = src.v;
              result.proposal = src.proposal;
              result.age = (src.age < 65535u ? src.age : 65535u) + 1u;
            }
            if (result.r.weight_sum > 0.0f) {
              bool valid = true;
              const float support = history_target<kLights>(S, src, psurf, result, valid);
              if (!valid) result.flags |= 2u;
              result.supporting = init.r.count + (support > 0.0f ? count : 0u);
            } else {
              result.supporting = 0u;
            }
            result.normalization = reuse_normalization(result.r, result.supporting);
          }
        }
      }
      resolve_estimate<kLights>(S, st, receiver, result, init, rays, visits, tests);
    }
    wave_count(restir_ray_count(counters), rays != 0u);
    store_reuse(A.temporal_out + 4ull * slot, result);
  }
  flush_stats(stats, 2, visits, tests);
}

// restir.slang:392-400 (SpatialCompatible): geometry alone.
__device__ __forceinline__ bool spatial_compatible(const ReuseSurface& receiver, uint32_t record_flags,
                                                   const ReuseSurface& source) {
  return (record_flags & 3u) == 1u && receiver.depth > 0.0f && source.depth > 0.0f &&
         receiver.material_id == source.material_id &&
         fabsf(receiver.depth - source.depth) <= 0.01f + 0.02f * maxf(receiver.depth, source.depth) &&
         dot3(receiver.normal, source.normal) >= 0.9f && dot3(receiver.shading_normal, source.shading_normal) >= 0.9f;
}

// The disk mode's random stream: neighbour offsets are drawn apart from the
// merges' stream (seeded with 0x369DEA0F), so both modes merge with the same
// random numbers.
constexpr uint32_t kSpatialDiskSeed = 0x5BD1E995u;
constexpr int kMaxSpatialSources = 5;

// restir.slang:402-491 (WriteSpatialReservoir): the four axis neighbours at
// the radius (restir.slang:408-410), or, in the disk mode, five neighbours
// drawn uniformly within the radius (a draw that lands on the pixel itself
// is no neighbour; draws may repeat a neighbour, whose candidates then count
// twice in M and in the supporting candidates alike).
template <bool kLights, bool kGltf = false>  // analytic lights or an environment image; glTF PBR receivers
__global__ void __launch_bounds__(kTraceBlock) k_restir_spatial(SceneView S, FrameConst F, RestirReuseArgs A,
                                                                uint32_t* counters, uint2* spill, int spill_cap,
                                                                unsigned long long* stats) {
  __shared__ uint2 ring[kRing * kTraceBlock];
  TraceStack st;
  st.init(ring, spill, spill_cap);
  const uint32_t pixels = F.width * F.height;
  const int radius = static_cast<int>(A.radius);
  const int sources = A.disk ? kMaxSpatialSources : 4;
  uint32_t visits = 0, tests = 0;
  for (uint32_t slot = gtid(); slot < pixels; slot += gstride()) {
    const ReuseRecord centre = load_reuse(A.temporal_out + 4ull * slot);
    ReuseRecord result = centre;
    uint32_t src_index[kMaxSpatialSources] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    uint32_t rays = 0;
    if ((centre.flags & 3u) == 1u) {
      const ReuseSurface receiver = load_reuse_surface<kGltf>(A.surf + kReuseSurfaceWords * slot);
      const int px = static_cast<int>(slot % F.width), py = static_cast<int>(slot / F.width);
      uint32_t disk_rng = seed_random(slot, F.sample_index ^ kSpatialDiskSeed);
      uint32_t accepted = 0;
      for (int i = 0; i < sources; ++i) {
        int ox, oy;
        if (A.disk) {
          const float u = next_random(disk_rng), v = next_random(disk_rng);
          const float rr = static_cast<float>(radius) * sqrtf(u);
          float sn, cs;
          sincos_turns(v, &sn, &cs);
          ox = static_cast<int>(floorf(rr * cs + 0.5f));
          oy = static_cast<int>(floorf(rr * sn + 0.5f));
          if (ox == 0 && oy == 0) continue;
        } else {
          ox = i == 0 ? -radius : i == 1 ? radius : 0;
          oy = i == 2 ? -radius : i == 3 ? radius : 0;
        }
        const int nx = px + ox, ny = py + oy;
        if (nx < 0 || ny < 0 || nx >= static_cast<int>(F.width) || ny >= static_cast<int>(F.height)) continue;
        const uint32_t index = static_cast<uint32_t>(ny) * F.width + static_cast<uint32_t>(nx);
        const uint32_t nflags = __float_as_uint(A.temporal_out[4ull * index + 1].w);
        if (spatial_compatible(receiver, nflags, load_reuse_surface<kGltf>(A.surf + kReuseSurfaceWords * index))) {
          src_index[i] = index;
          ++accepted;
        }
      }
      if (accepted != 0u) {
        result.r = empty_reservoir();
        bool selected = false;
        if (!merge_reevaluated(result.r, centre.r, centre.normalization, centre.r.selected_target, 0xFFFFFFFFu, 0.0f,
                               selected))
          result.flags |= 2u;
        uint32_t rng = seed_random(slot, F.sample_index ^ 0x369DEA0Fu);
        for (int i = 0; i < sources; ++i) {
          if (src_index[i] == 0xFFFFFFFFu) continue;
          const ReuseRecord src = load_reuse(A.temporal_out + 4ull * src_index[i]);
          V3 received = V3{0.0f, 0.0f, 0.0f};
          if (src.r.weight_sum > 0.0f) received = reuse_contribution_uv<kLights>(S, receiver, src.r.sample, src.u, src.v);
          if (!finite3(received) || !merge_reevaluated(result.r, src.r, src.normalization, max3(received),
                                                       A.max_history, next_random(rng), selected)) {
            result.flags |= 2u;
            break;
          }
          if (selected) {
            result.u = src.u;
            result.v = src.v;
            result.proposal = src.proposal;
            result.age = src.age;
          }
        }
        result.supporting = 0u;
        result.contribution = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if ((result.flags & 2u) == 0u && result.r.weight_sum > 0.0f) {
          result.supporting = centre.r.count;
          for (int i = 0; i < sources; ++i) {
            if (src_index[i] == 0xFFFFFFFFu) continue;
            const ReuseRecord src = load_reuse(A.temporal_out + 4ull * src_index[i]);
            bool valid = true;
            const float support =
                history_target<kLights>(S, src, load_reuse_surface<kGltf>(A.surf + kReuseSurfaceWords * src_index[i]), result, valid);
            if (!valid) result.flags |= 2u;
            if (support > 0.0f) result.supporting += src.r.count < A.max_history ? src.r.count : A.max_history;
          }
        }
        result.normalization = reuse_normalization(result.r, result.supporting);
        resolve_estimate<kLights>(S, st, receiver, result, centre, rays, visits, tests);
      }
    }
    wave_count(restir_ray_count(counters), rays != 0u);
    store_reuse(A.spatial_out + 4ull * slot, result);
    A.spatial_sources[2ull * slot] = make_uint4(src_index[0], src_index[1], src_index[2], src_index[3]);
    A.spatial_sources[2ull * slot + 1] = make_uint4(src_index[4], 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
  }
  flush_stats(stats, 2, visits, tests);
}

// A cache query for a surface reached along `dir`: position in the cache's
// unit cube, direction towards the previous vertex, shading normal, material.
__device__ __forceinline__ void cache_query(const FrameConst& F, const Surface& sf, V3 dir, float4* out) {
  const float px = (sf.position.x - F.cache_lo[0]) * F.cache_scale[0];
  const float py = (sf.position.y - F.cache_lo[1]) * F.cache_scale[1];
  const float pz = (sf.position.z - F.cache_lo[2]) * F.cache_scale[2];
  const V3 n = sf.shading_normal;
  out[0] = make_float4(px, py, pz, -dir.x);
  out[1] = make_float4(-dir.y, -dir.z, n.x, n.y);
  out[2] = make_float4(n.z, sf.roughness, sf.base_color.x, sf.base_color.y);
  out[3] = make_float4(sf.base_color.z, sf.metallic, 0.0f, 0.0f);
}

// Under BSDF sampling with several parallel lights, ParallelLights
// (transport.slang:139-146) at the shading lane itself: each parallel light's
// ShadowedLight (:82-115) with a shadow ray the lane walks at once, summed in
// light order (a blocked or non-finite term adds nothing), the rays counted
// in parallel_ray_count.
template <bool kCov>
struct ParallelWalk {
  static constexpr bool kEnabled = true;
  TraceStack* st;
  uint32_t* rays;
  __device__ V3 operator()(const SceneView& S, const Surface& sf, const V3& ft, const V3& fb, const V3& wo) const {
    const V3 n = sf.shading_normal;
    const V3 origin = offset_ray_origin(sf.position, sf.normal);
    V3 sum{0.0f, 0.0f, 0.0f};
    uint32_t traced = 0;
    for (uint32_t a = 0; a < S.analytic_count; ++a) {
      const AnalyticRecord& L = S.analytic[a];
      if (L.shape != kParallelLight) continue;  // ParallelLight gives NoLight for the others
      const V3 dir{L.axis_u[0], L.axis_u[1], L.axis_u[2]}, rad{L.radiance[0], L.radiance[1], L.radiance[2]};
      if (max3(rad) <= 0.0f || dot3(dir, sf.normal) <= 0.0f) continue;
      const V3 wi{dot3(dir, ft), dot3(dir, fb), dot3(dir, n)};
      if (wi.z <= 0.0f) continue;
      const V3 f = bsdf_evaluate(sf.bsdf, wo, wi);
      if (max3(f) <= 0.0f) continue;
      ++traced;
      uint32_t visits = 0, tests = 0;
      const float t = walk_transmittance(S.nodes, S.tris, *st, Ray{origin, 0.0f, dir, kFloatMax}, visits, tests,
                                         coverage_of<kCov>(S));
      if (t <= 0.0f) continue;
      // f * radiance * (transmittance * weight / pdf), weight and pdf 1
      const V3 c = f * rad * ((t * 1.0f) / 1.0f);
      if (finite3(c)) sum = sum + c;
    }
    if (traced != 0) atomicAdd(rays, traced);
    return sum;
  }
};

// wavefront.slang:264-282: one scattering event per queued hit. The light a
// shadow ray of the previous bounce found is added first. In radiance cache
// frames a path other than a training path ends at bounce F.cache_depth after
// its emission term, leaving a cache query with its throughput, and a
// training path records each vertex (query, throughput on arrival, radiance
// including the vertex's emission).
// kParallel: BSDF sampling walks several parallel lights (ParallelWalk), with
// traversal stacks in LDS.
template <bool kAnalytic, bool kCov, bool kParallel = false,
          bool kGltf = false>  // analytic lights (extras); covered surfaces; glTF PBR surfaces
__global__ void __launch_bounds__(kShadeBlock) k_shade(SceneView S, FrameConst F, Queues q, uint32_t* counters,
                                                       uint32_t bounce, uint32_t link, uint2* spill, int spill_cap) {
  using Walk = std::conditional_t<kParallel, ParallelWalk<kCov>, NoParallelWalk>;
  Walk walk{};
  __shared__ uint2 ring[kParallel ? kRing * kTraceBlock : 1];
  TraceStack st;
  if constexpr (kParallel) {
    st.init(ring, spill, spill_cap);
    walk.st = &st;
    walk.rays = parallel_ray_count(counters);
  }
  const uint32_t count = *hit_count(counters, bounce);
  for (uint32_t i = gtid(); i < count; i += gstride()) {
    const float4 o4 = q.hit_o[i], d4 = q.hit_d[i];
    const uint4 h4 = q.hit_h[i];
    const uint32_t slot = __float_as_uint(o4.w);
    PathRegs p;
    p.o = xyz(o4);
    p.d = xyz(d4);
    p.depth = bounce;
    const float4 r4 = q.rad[slot];
    p.radiance = xyz(r4);
    p.rng = __float_as_uint(r4.w);
    if (bounce == 0) {
      p.throughput = V3{1.0f, 1.0f, 1.0f};
      p.last_pdf = 0.0f;
    } else {
      const float4 t4 = q.thr[slot];
      p.throughput = xyz(t4);
      p.last_pdf = t4.w;
      const float4 n4 = q.nee[slot];
      if (n4.w != 0.0f) {
        p.radiance = p.radiance + xyz(n4);
        q.nee[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
    }
    bool train = false, query = false;
    uint32_t path = 0;
    if (F.cache_depth != 0) {
      const uint32_t key = slot + F.train_offset;
      train = key % F.train_stride == 0;
      path = key / F.train_stride;
      query = !train && bounce == F.cache_depth;
    }
    const V3 arrival = p.throughput, dir = p.d;
    Scatter s;
    Surface sf;
    ShadowFactors factors;
    if (query) {
      sf = reconstruct_hit<kAnalytic, kGltf>(S, __float_as_uint(d4.w), h4.x, __uint_as_float(h4.z),
                                             __uint_as_float(h4.w), p.d);
      add_emission(S, F, p, sf, h4.x, h4.y);
      s.continues = false;
      s.shadow = false;
    } else {
      s = scatter_path<kAnalytic, kCov, Walk, kGltf>(S, F, p, __float_as_uint(d4.w), h4.x, h4.y, __uint_as_float(h4.z),
                                              __uint_as_float(h4.w), train ? &sf : nullptr, slot,
                                              kCov ? &factors : nullptr, walk);
    }
    const uint32_t kq = wave_append(cache_query_count(counters), query);
    if (query) {
      cache_query(F, sf, dir, q.cq + 4ull * kq);
      q.cq_thr[kq] = make_float4(p.throughput.x, p.throughput.y, p.throughput.z, __uint_as_float(slot));
    }
    if (train && bounce < kCacheRecordDepth && path < F.train_capacity) {
      float4* rec = q.crec + 6ull * (static_cast<size_t>(bounce) * F.train_capacity + path);
      cache_query(F, sf, dir, rec);
      rec[4] = make_float4(arrival.x, arrival.y, arrival.z, 1.0f);
      rec[5] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 0.0f);
      q.crec_n[path] = bounce + 1;
    }
    const uint32_t ks = wave_append(shadow_count(counters, bounce), s.shadow);
    if (s.shadow) {
      q.sh_o[ks] = make_float4(s.shadow_ray.o.x, s.shadow_ray.o.y, s.shadow_ray.o.z, __uint_as_float(slot));
      q.sh_d[ks] = make_float4(s.shadow_ray.d.x, s.shadow_ray.d.y, s.shadow_ray.d.z, s.shadow_ray.tmax);
      // w: the tail walks this ray with its path (1) or on its own (0)
      q.sh_c[ks] = make_float4(s.delta.x, s.delta.y, s.delta.z, __uint_as_float(link && s.continues ? 1u : 0u));
      if constexpr (kCov) {
        q.sh_f[ks] = make_float4(factors.fl.x, factors.fl.y, factors.fl.z, factors.weight);
        q.sh_t[ks] = make_float4(factors.thr.x, factors.thr.y, factors.thr.z, factors.pdf);
      }
    }
    const uint32_t kr = wave_append(ray_count(counters, bounce + 1), s.continues);
    if (s.continues) {
      q.thr[slot] = make_float4(p.throughput.x, p.throughput.y, p.throughput.z, p.last_pdf);
      q.rad[slot] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, __uint_as_float(p.rng));
      q.ray_o[kr] = make_float4(p.o.x, p.o.y, p.o.z, __uint_as_float(slot));
      // before the tail, w links the path's shadow ray of this bounce (the
      // tail's rays need no tmax), else the extension ray's tmax
      q.ray_d[kr] = make_float4(p.d.x, p.d.y, p.d.z,
                                link ? __uint_as_float(s.shadow ? ks : kNoLink) : kFloatMax);
    } else {
      q.rad[slot] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 1.0f);
    }
  }
}

// Training samples of a radiance cache frame, after the resolve inputs are
// final: a training path's sample value L (radiance, pending direct light,
// pending environment, summed as the resolve sums them) gives each recorded
// vertex the target (L - R) / T, R the radiance summed up to and including
// the vertex's emission and T the throughput on arrival. Vertices whose
// throughput has a component below kCacheMinThroughput, and non-finite
// targets, are skipped.
__global__ void __launch_bounds__(256) k_cache_samples(FrameConst F, Queues q, uint32_t* counters,
                                                       uint32_t capacity) {
  const uint32_t P = F.width * F.height;
  for (uint32_t path = gtid(); path < F.train_capacity; path += gstride()) {
    const uint32_t key = path * F.train_stride;
    const uint32_t n = q.crec_n[path];
    V3 L = V3{0.0f, 0.0f, 0.0f};
    if (n != 0 && key >= F.train_offset && key - F.train_offset < P) {
      const uint32_t slot = key - F.train_offset;
      const float4 s = q.rad[slot], ne = q.nee[slot], e = q.esc[slot];
      L = xyz(s);
      if (ne.w != 0.0f) L = L + xyz(ne);
      if (e.w != 0.0f) L = L + xyz(e);
      if (e.w == 2.0f) L = L + xyz(q.esc2[slot]);
    }
    for (uint32_t b = 0; b < kCacheRecordDepth; ++b) {
      bool ok = b < n;
      float4 rec[6];
      V3 target = V3{0.0f, 0.0f, 0.0f};
      if (ok) {
        const float4* r = q.crec + 6ull * (static_cast<size_t>(b) * F.train_capacity + path);
#pragma unroll
        for (int k = 0; k < 6; ++k) rec[k] = r[k];
        const V3 T = xyz(rec[4]), R = xyz(rec[5]);
        ok = T.x >= kCacheMinThroughput && T.y >= kCacheMinThroughput && T.z >= kCacheMinThroughput;
        if (ok) {
          target = V3{(L.x - R.x) / T.x, (L.y - R.y) / T.y, (L.z - R.z) / T.z};
          ok = finite3(target);
        }
      }
      const uint32_t k = wave_append(cache_sample_count(counters), ok);
      if (ok && k < capacity) {
        float4* dst = reinterpret_cast<float4*>(q.samples + static_cast<size_t>(k) * kSampleFloats);
#pragma unroll
        for (int j = 0; j < 4; ++j) dst[j] = rec[j];
        dst[4] = make_float4(target.x, target.y, target.z, 0.0f);
      }
    }
  }
}

// wavefront.slang:230-262 and ShadowTransmittance (scene.slang:445-460) in one
// grid: items [0, n_ext) are the extension rays of bounce b (a hit is queued
// for shading, a miss leaves its environment term pending, transport.slang:
// 436-448), items [n_ext, n_ext + n_shadow) the shadow rays of bounce b - 1
// (an unblocked ray leaves its light pending).
template <bool kDynamic, bool kCov>  // kCov: the scene has covered surfaces
__global__ void __launch_bounds__(kTraceBlock) k_trace(SceneView S, Queues q, uint32_t* counters, uint2* spill,
                                                       int spill_cap, uint32_t bounce, uint32_t kinds,
                                                       unsigned long long* stats) {
  __shared__ uint2 ring[kRing * kTraceBlock];
  TraceStack st;
  st.init(ring, spill, spill_cap);
  const uint32_t n_ext = (kinds & kTraceExtension) ? *ray_count(counters, bounce) : 0u;
  const uint32_t n_sh = (kinds & kTraceShadow) ? *shadow_count(counters, bounce - 1) : 0u;
  const bool lights = scene_has_extras(S);  // escapes go to k_escape
  __shared__ unsigned long long counts[4];
  BlockStats<2> bs;
  bs.init(counts);
  walks<kDynamic>(
      S, st, n_ext + n_sh,
      [&](uint32_t i, auto& w) {
        const bool ext = i < n_ext;
        const uint32_t j = ext ? i : i - n_ext;
        const float4 o4 = ext ? q.ray_o[j] : q.sh_o[j];
        const float4 d4 = ext ? q.ray_d[j] : q.sh_d[j];
        w.start(Ray{xyz(o4), 0.0f, xyz(d4), d4.w}, !ext);
      },
      [&](uint32_t i, const auto& w) {
        const bool ext = i < n_ext;
        const uint32_t j = ext ? i : i - n_ext;
        const uint32_t slot = __float_as_uint(ext ? q.ray_o[j].w : q.sh_o[j].w);
        const bool hit = w.hit();
        bs.add(ext ? 0 : 1, w.visits, w.tests);
        const uint32_t k = wave_append(hit_count(counters, bounce), ext && hit);
        const uint32_t m = wave_append(escape_count(counters, bounce), ext && !hit && lights);
        if (ext && hit) {
          push_hit(q, k, slot, w.r.o, w.r.d, w.record(S.tris));
        } else if (ext) {
          if (lights) {
            q.esc_o[m] = make_float4(w.r.o.x, w.r.o.y, w.r.o.z, __uint_as_float(slot));
            q.esc_d[m] = make_float4(w.r.d.x, w.r.d.y, w.r.d.z, 0.0f);
          } else {
            const V3 e = xyz(q.thr[slot]) * V3{S.environment[0], S.environment[1], S.environment[2]};
            q.esc[slot] = make_float4(e.x, e.y, e.z, 1.0f);
          }
          float4 r4 = q.rad[slot];
          r4.w = 1.0f;
          q.rad[slot] = r4;
        } else {
          float4 nee;
          if (shadow_light(q, j, w, nee)) q.nee[slot] = nee;
        }
      },
      coverage_of<kCov>(S));
  bs.flush(stats, 1);
}

// transport.slang:235-247 (MissPath) for the escaped extension rays of a
// round in a scene with analytic lights or an environment image: each
// slot's pending environment term (the constant plus the image, weighed
// against light sampling) and the analytic lights the ray crossed, kept apart
// so radiance is summed in the reference's order. Kept out of k_trace, whose
// registers then stay those of the scenes without lights.
__global__ void __launch_bounds__(256) k_escape(SceneView S, Queues q, uint32_t* counters, uint32_t bounce,
                                                uint32_t weigh) {
  const uint32_t n = *escape_count(counters, bounce);
  for (uint32_t i = gtid(); i < n; i += gstride()) {
    const float4 o4 = q.esc_o[i], d4 = q.esc_d[i];
    const uint32_t slot = __float_as_uint(o4.w);
    const float4 t4 = q.thr[slot];
    V3 e, crossed;
    escape_terms<true>(S, weigh != 0, xyz(o4), xyz(d4), xyz(t4), t4.w, e, crossed);
    const bool more = S.analytic_count != 0;
    q.esc[slot] = make_float4(e.x, e.y, e.z, more ? 2.0f : 1.0f);
    if (more) q.esc2[slot] = make_float4(crossed.x, crossed.y, crossed.z, 0.0f);
  }
}

// wavefront.slang:287-309: the paths still queued after the recorded rounds,
// each followed to its end in one invocation. A path has a pair of lanes:
// the even lane holds the path and walks its extension rays, the odd lane
// walks the shadow ray its previous bounce sent, both at once, so a bounce
// costs the longer of its two walks rather than their sum. The direct light
// of bounce b is added before anything bounce b + 1 adds, as in a serial
// walk, so radiance is summed in the reference's order.
__device__ __forceinline__ float pair_float(float v) { return __shfl_xor(v, 1); }

// a radiance cache frame (training vertices recorded); analytic lights
// (extras); covered surfaces
// (extras); covered surfaces; BSDF sampling walks several parallel lights
// (ParallelWalk, on the stack the lane's walks use, which is free while it
// shades)
template <bool kCache, bool kAnalytic, bool kCov, bool kParallel = false, bool kGltf = false>
__global__ void __launch_bounds__(kTraceBlock) k_tail(SceneView S, FrameConst F, Queues q, uint32_t* counters,
                                                      uint2* spill, int spill_cap, uint32_t bounce,
                                                      unsigned long long* stats) {
  __shared__ uint2 ring[kRing * kTraceBlock];
  TraceStack st;
  st.init(ring, spill, spill_cap);
  using Walk = std::conditional_t<kParallel, ParallelWalk<kCov>, NoParallelWalk>;
  Walk walk{};
  if constexpr (kParallel) {
    walk.st = &st;
    walk.rays = parallel_ray_count(counters);
  }
  const uint32_t count = *ray_count(counters, bounce);
  const uint32_t n_shadow = *shadow_count(counters, bounce - 1);  // the last round's, linked or on their own
  const bool path_lane = (threadIdx.x & 1u) == 0;
  uint32_t visits = 0, tests = 0;
  // both lanes of a pair take the same items (the grid's workgroups hold
  // whole pairs), so the pair enters and leaves the loops together: items
  // [0, count) are paths, then the last round's shadow rays
  for (uint32_t i = gtid() >> 1; i < count + n_shadow; i += gstride() >> 1) {
    if (i >= count) {
      // a shadow ray of a path that ended at the last round: the odd lane
      // walks it, and an unblocked ray leaves its light pending (as k_trace)
      const uint32_t j = i - count;
      const float4 c4 = q.sh_c[j];
      const bool alone = __float_as_uint(c4.w) == 0u;
      WalkT<CoverageOf<kCov>> w;
      w.cov = coverage_of<kCov>(S);
      uint32_t sslot = 0;
      if (!path_lane && alone) {
        const float4 o4 = q.sh_o[j], d4 = q.sh_d[j];
        sslot = __float_as_uint(o4.w);
        w.start(Ray{xyz(o4), 0.0f, xyz(d4), d4.w}, true);
      } else {
        w.start(Ray{V3{0.0f, 0.0f, 0.0f}, 0.0f, V3{0.0f, 0.0f, 1.0f}, 0.0f}, true);
        w.cur = kWalkDone;
      }
      st.clear();
      w.run(S.nodes, S.tris, st);
      visits += w.visits;
      tests += w.tests;
      if constexpr (kCov) {
        float4 nee;
        if (!path_lane && alone && shadow_light(q, j, w, nee)) q.nee[sslot] = nee;
      } else {
        if (!path_lane && alone && !w.hit()) q.nee[sslot] = make_float4(c4.x, c4.y, c4.z, 1.0f);
      }
      continue;
    }
    PathRegs p{};
    p.depth = bounce;
    uint32_t slot = 0;
    if (path_lane) {
      const float4 o4 = q.ray_o[i], d4 = q.ray_d[i];
      slot = __float_as_uint(o4.w);
      const float4 t4 = q.thr[slot], r4 = q.rad[slot], n4 = q.nee[slot];
      p.o = xyz(o4);
      p.d = xyz(d4);
      p.throughput = xyz(t4);
      p.last_pdf = t4.w;
      p.radiance = xyz(r4);
      if (n4.w != 0.0f) {
        p.radiance = p.radiance + xyz(n4);
        q.nee[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
      p.rng = __float_as_uint(r4.w);
    }
    // radiance cache frames: only training paths reach the tail (the others
    // end at the cache depth, inside the rounds); their vertices are recorded
    // as the shading kernel records them
    bool train = false;
    uint32_t path = 0;
    if (kCache && path_lane) {
      const uint32_t key = slot + F.train_offset;
      train = key % F.train_stride == 0;
      path = key / F.train_stride;
    }
    bool extend = path_lane;  // the path lane has an extension ray to walk
    bool shadow = false;      // a shadow ray of the previous bounce is pending
    Ray sray{};               // that ray (sent to the odd lane)
    V3 delta{};               // its contribution (kept in the path lane)
    ShadowFactors sf4{};      // with covered surfaces: its factors
    if (path_lane) {
      // the shadow ray the last round's shading linked to this path
      const uint32_t link = __float_as_uint(q.ray_d[i].w);
      if (link != kNoLink) {
        const float4 o4 = q.sh_o[link], d4 = q.sh_d[link], c4 = q.sh_c[link];
        shadow = true;
        sray = Ray{xyz(o4), 0.0f, xyz(d4), d4.w};
        delta = xyz(c4);
        if constexpr (kCov) {
          const float4 f4 = q.sh_f[link], t4 = q.sh_t[link];
          sf4 = ShadowFactors{xyz(f4), f4.w, t4.w, xyz(t4)};
        }
      }
    }
    while (true) {
      // the pair's state, as the path lane holds it
      const bool pair_extend = __shfl_xor(static_cast<int>(extend), 1) != 0 || extend;
      const bool pair_shadow = __shfl_xor(static_cast<int>(shadow), 1) != 0 || shadow;
      if (!pair_extend && !pair_shadow) break;
      // the odd lane receives the shadow ray
      const Ray sr{V3{pair_float(sray.o.x), pair_float(sray.o.y), pair_float(sray.o.z)}, pair_float(sray.tmin),
                   V3{pair_float(sray.d.x), pair_float(sray.d.y), pair_float(sray.d.z)}, pair_float(sray.tmax)};
      WalkT<CoverageOf<kCov>> w;
      w.cov = coverage_of<kCov>(S);
      if (path_lane) {
        w.start(Ray{p.o, 0.0f, p.d, kFloatMax}, false);
        if (!extend) w.cur = kWalkDone;
      } else {
        w.start(sr, true);
        if (!pair_shadow) w.cur = kWalkDone;
      }
      st.clear();
      w.run(S.nodes, S.tris, st);
      visits += w.visits;
      tests += w.tests;
      // the odd lane's answer: blocked, or with covered surfaces its
      // transmittance (0: blocked)
      if constexpr (kCov) {
        const float transmittance = pair_float(w.transmittance());
        if (path_lane && shadow && transmittance > 0.0f)
          p.radiance = p.radiance + shadowed_delta(sf4.fl, sf4.weight, sf4.pdf, sf4.thr, transmittance);
      } else {
        const bool blocked = __shfl_xor(static_cast<int>(w.hit()), 1) != 0;
        if (path_lane && shadow && !blocked) p.radiance = p.radiance + delta;
      }
      shadow = false;
      Scatter sc{};
      ShadowFactors sc4;
      bool hit = false;
      if (path_lane && extend) {
        hit = w.hit();
        if (!hit) {
          V3 e, crossed;
          escape_terms<kAnalytic>(S, F.light_sampling != 0, p.o, p.d, p.throughput, p.last_pdf, e, crossed);
          p.radiance = p.radiance + e;
          if (kAnalytic && S.analytic_count != 0) p.radiance = p.radiance + crossed;
        } else {
          const HitRecord h = w.record(S.tris);
          const V3 arrival = p.throughput, dir = p.d;
          const uint32_t b = p.depth;
          Surface sf;
          sc = scatter_path<kAnalytic, kCov, Walk, kGltf>(S, F, p, h.tri, h.instance, h.primitive, h.u, h.v,
                                                   kCache && train ? &sf : nullptr, slot, kCov ? &sc4 : nullptr, walk);
          if (kCache && train && b < kCacheRecordDepth && path < F.train_capacity) {
            float4* rec = q.crec + 6ull * (static_cast<size_t>(b) * F.train_capacity + path);
            cache_query(F, sf, dir, rec);
            rec[4] = make_float4(arrival.x, arrival.y, arrival.z, 1.0f);
            rec[5] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 0.0f);
            q.crec_n[path] = b + 1;
          }
        }
      }
      const bool shaded = path_lane && extend && hit;
      wave_count_each(shadow_count(counters, p.depth - (sc.continues ? 1u : 0u)), shaded && sc.shadow);
      wave_count_each(ray_count(counters, p.depth), shaded && sc.continues);
      if (shaded && sc.shadow) {
        shadow = true;
        sray = sc.shadow_ray;
        delta = sc.delta;
        if constexpr (kCov) sf4 = sc4;
      }
      extend = shaded && sc.continues;
    }
    if (path_lane) q.rad[slot] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 1.0f);
  }
  flush_stats(stats, 3, visits, tests);
}

// A whole path in one lane pair, from its camera ray to its end: the path
// lane walks the camera ray and then each extension ray and shades every
// hit (ScatterPath, transport.slang:164-230); the odd lane walks the shadow
// ray of the previous bounce beside it, as the tail does. Path i of the
// data window is pixel (i mod 32) of its 8 x 4 tile i / 32, so a wave's 32
// pairs start from neighbouring pixels. A camera miss leaves no sample, as
// wavefront.slang:206-225 does; counts are those of the rounds (each ray at
// its bounce of departure), and the sample the resolve reads is complete,
// with no light or environment term pending. Both lanes of a pair call this
// with the same i.
template <bool kAnalytic, bool kCov>
__device__ __forceinline__ void camera_path(const SceneView& S, const FrameConst& F, const Queues& q,
                                            uint32_t* counters, TraceStack& st, uint32_t i, uint32_t& visits,
                                            uint32_t& tests) {
  const bool path_lane = (threadIdx.x & 1u) == 0;
  const uint32_t rw = F.rect[2], rh = F.rect[3];
  const uint32_t tiles_x = (rw + 7) / 8, tile = i >> 5, within = i & 31u;
  const uint32_t lx = (tile % tiles_x) * 8 + (within & 7u), ly = (tile / tiles_x) * 4 + (within >> 3);
  if (lx >= rw || ly >= rh) return;
  const uint32_t x = F.rect[0] + lx, y = F.rect[1] + ly, slot = y * F.width + x;
  PathRegs p{};
  Ray camera{};
  if (path_lane) {
    uint32_t rng;
    camera = start_path(F, x, y, frame_sample_index(F), rng);
    p.o = camera.o;
    p.d = camera.d;
    p.depth = 0;
    p.throughput = V3{1.0f, 1.0f, 1.0f};
    p.last_pdf = 0.0f;
    p.radiance = V3{0.0f, 0.0f, 0.0f};
    p.rng = rng;
  }
  bool first = true, missed = false;
  bool extend = path_lane;  // the path lane has a ray to walk
  bool shadow = false;      // a shadow ray of the previous bounce is pending
  Ray sray{};
  V3 delta{};
  ShadowFactors sf4{};
  while (true) {
    const bool pair_extend = __shfl_xor(static_cast<int>(extend), 1) != 0 || extend;
    const bool pair_shadow = __shfl_xor(static_cast<int>(shadow), 1) != 0 || shadow;
    if (!pair_extend && !pair_shadow) break;
    const Ray sr{V3{pair_float(sray.o.x), pair_float(sray.o.y), pair_float(sray.o.z)}, pair_float(sray.tmin),
                 V3{pair_float(sray.d.x), pair_float(sray.d.y), pair_float(sray.d.z)}, pair_float(sray.tmax)};
    WalkT<CoverageOf<kCov>> w;
    w.cov = coverage_of<kCov>(S);
    if (path_lane) {
      w.start(first ? camera : Ray{p.o, 0.0f, p.d, kFloatMax}, false);
      if (!extend) w.cur = kWalkDone;
    } else {
      w.start(sr, true);
      if (!pair_shadow) w.cur = kWalkDone;
    }
    st.clear();
    w.run(S.nodes, S.tris, st);
    visits += w.visits;
    tests += w.tests;
    if constexpr (kCov) {
      const float transmittance = pair_float(w.transmittance());
      if (path_lane && shadow && transmittance > 0.0f)
        p.radiance = p.radiance + shadowed_delta(sf4.fl, sf4.weight, sf4.pdf, sf4.thr, transmittance);
    } else {
      const bool blocked = __shfl_xor(static_cast<int>(w.hit()), 1) != 0;
      if (path_lane && shadow && !blocked) p.radiance = p.radiance + delta;
    }
    shadow = false;
    Scatter sc{};
    ShadowFactors sc4;
    bool hit = false;
    if (path_lane && extend) {
      hit = w.hit();
      if (!hit) {
        if (first) {
          missed = true;
        } else {
          V3 e, crossed;
          escape_terms<kAnalytic>(S, F.light_sampling != 0, p.o, p.d, p.throughput, p.last_pdf, e, crossed);
          p.radiance = p.radiance + e;
          if (kAnalytic && S.analytic_count != 0) p.radiance = p.radiance + crossed;
        }
      } else {
        const HitRecord h = w.record(S.tris);
        sc = scatter_path<kAnalytic, kCov>(S, F, p, h.tri, h.instance, h.primitive, h.u, h.v, nullptr, slot,
                                           kCov ? &sc4 : nullptr);
      }
    }
    const bool shaded = path_lane && extend && hit;
    wave_count_each(shadow_count(counters, p.depth - (sc.continues ? 1u : 0u)), shaded && sc.shadow);
    wave_count_each(ray_count(counters, p.depth), shaded && sc.continues);
    if (shaded && sc.shadow) {
      shadow = true;
      sray = sc.shadow_ray;
      delta = sc.delta;
      if constexpr (kCov) sf4 = sc4;
    }
    extend = shaded && sc.continues;
    first = false;
  }
  if (path_lane) {
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    q.rad[slot] = missed ? zero : make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 1.0f);
    q.nee[slot] = zero;
    q.esc[slot] = zero;
  }
}

// Paths of a data window of 8 x 4 tiles.
__host__ __device__ inline uint32_t window_paths(uint32_t rw, uint32_t rh) {
  return ((rw + 7) / 8) * ((rh + 3) / 4) * 32u;
}

// The path megakernel of one view: every path of the data window from its
// camera ray to its end (camera_path), the pairs striding over the paths.
template <bool kAnalytic, bool kCov>
__global__ void __launch_bounds__(kTraceBlock) k_paths(SceneView S, FrameConst F, Queues q, uint32_t* counters,
                                                       uint2* spill, int spill_cap, unsigned long long* stats) {
  __shared__ uint2 ring[kRing * kTraceBlock];
  TraceStack st;
  st.init(ring, spill, spill_cap);
  const uint32_t rw = F.rect[2], rh = F.rect[3];
  if (gtid() == 0) *ray_count(counters, 0) = rw * rh;
  const uint32_t n = window_paths(rw, rh);
  uint32_t visits = 0, tests = 0;
  for (uint32_t i = gtid() >> 1; i < n; i += gstride() >> 1)
    camera_path<kAnalytic, kCov>(S, F, q, counters, st, i, visits, tests);
  flush_stats(stats, 3, visits, tests);
}

// The path megakernel of a batch of views fed by one device work queue, cut
// for the device's eight XCDs (each with its own L2): the batch's chunks of
// kPathChunk paths (views in table order) split into eight contiguous
// ranges, one counter each (next_chunk[0..7]). Workgroups are dispatched to
// the XCDs in turn (XCD = workgroup index mod 8), so a wave starts on its
// own XCD's range, which holds a few neighbouring views, and so their trees
// stay in that L2; once that range is drained it helps the next ranges in
// turn. Each grab is one atomic; the wave finds the chunk's view in the view
// table (each with its first chunk) and runs its paths with that view's
// scene, frame and buffers. The grid drains when every range has.
template <bool kAnalytic, bool kCov>
__global__ void __launch_bounds__(kTraceBlock) k_paths_views(const PathView* __restrict__ views, uint32_t view_count,
                                                             uint32_t* next_chunk, uint32_t total_chunks,
                                                             uint2* spill, int spill_cap) {
  __shared__ uint2 ring[kRing * kTraceBlock];
  TraceStack st;
  st.init(ring, spill, spill_cap);
  const uint32_t lane = __lane_id();
  uint32_t part = blockIdx.x % kXcds, drained = 0;
  while (drained < kXcds) {
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(next_chunk + part, 1u);
    k = __shfl(k, 0);
    const uint32_t first = static_cast<uint32_t>((uint64_t{total_chunks} * part) / kXcds);
    const uint32_t end = static_cast<uint32_t>((uint64_t{total_chunks} * (part + 1)) / kXcds);
    if (first + k >= end) {
      part = part + 1 == kXcds ? 0 : part + 1;
      ++drained;
      continue;
    }
    const uint32_t c = first + k;
    // the last view whose first chunk is at most c (binary search)
    uint32_t lo = 0, hi = view_count;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (views[mid].first_chunk <= c) lo = mid; else hi = mid;
    }
    const PathView& V = views[lo];
    const uint32_t base = (c - V.first_chunk) * kPathChunk;
    uint32_t visits = 0, tests = 0;
    for (uint32_t j = lane >> 1; j < kPathChunk; j += 32) {
      const uint32_t i = base + j;
      if (i >= V.paths) break;
      if (i == 0 && (lane & 1u) == 0) *ray_count(V.counters, 0) = V.F.rect[2] * V.F.rect[3];
      camera_path<kAnalytic, kCov>(V.S, V.F, V.q, V.counters, st, i, visits, tests);
    }
    flush_stats(V.stats, 3, visits, tests);
  }
}

// path_trace.slang:106-139 (kResolvePass): each slot's sample (its radiance,
// then the pending direct light, then the pending environment) added to the
// pixel's accumulation, and the mean written, a missed sample counting as the
// background; pixels no sample has hit keep the clear colour.
__global__ void __launch_bounds__(256) k_resolve(FrameConst F, Queues q, uint32_t add_arg) {
  const uint32_t rw = F.rect[2], total = F.rect[2] * F.rect[3];
  const uint32_t accumulated = frame_accumulated(F);
  const uint32_t add_sample = F.vars ? F.vars[2] : add_arg;
  for (uint32_t i = gtid(); i < total; i += gstride()) {
    const uint32_t x = F.rect[0] + i % rw, y = F.rect[1] + i / rw;
    const uint32_t slot = y * F.width + x;
    float4 sum = accumulated != 0 ? q.accum[slot] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (add_sample) {
      const float4 s = q.rad[slot], n = q.nee[slot], e = q.esc[slot];
      V3 v = xyz(s);
      if (n.w != 0.0f) v = v + xyz(n);
      if (e.w != 0.0f) v = v + xyz(e);
      if (e.w == 2.0f) v = v + xyz(q.esc2[slot]);
      sum.x = sum.x + v.x;
      sum.y = sum.y + v.y;
      sum.z = sum.z + v.z;
      sum.w = sum.w + s.w;
      q.accum[slot] = sum;
    }
    const float count = static_cast<float>(accumulated + add_sample);
    if (sum.w == 0.0f || count == 0.0f) {
      q.color[slot] = make_float4(F.background[0], F.background[1], F.background[2], F.background[3]);
      continue;
    }
    const float miss = count - sum.w;
    q.color[slot] = make_float4((sum.x + miss * F.background[0]) / count, (sum.y + miss * F.background[1]) / count,
                                (sum.z + miss * F.background[2]) / count, (sum.w + miss * F.background[3]) / count);
  }
}

// path_trace.slang:69-97 (CameraPass): the pixel centre's closest hit as
// depth and barycentric weights or a surface diagnostic. Misses keep the
// clears (colour = background, depth = 1).
__global__ void __launch_bounds__(kTraceBlock) k_camera_pass(SceneView S, FrameConst F, Queues q, float* depth,
                                                             uint32_t write_color, uint2* spill, int spill_cap) {
  __shared__ uint2 ring[kRing * kTraceBlock];
  TraceStack st;
  st.init(ring, spill, spill_cap);
  const uint32_t rw = F.rect[2], total = F.rect[2] * F.rect[3];
  uint32_t visits = 0, tests = 0;
  for (uint32_t i = gtid(); i < total; i += gstride()) {
    const uint32_t x = F.rect[0] + i % rw, y = F.rect[1] + i / rw;
    const uint32_t slot = y * F.width + x;
    const float fx = static_cast<float>(x) + 0.5f, fy = static_cast<float>(y) + 0.5f;
    const Ray r = camera_ray(F.clip_to_world, fx * F.ndc_from_pixel[0] + F.ndc_from_pixel[2],
                             fy * F.ndc_from_pixel[1] + F.ndc_from_pixel[3]);
    HitRecord h;
    if (!walk_closest(S.nodes, S.tris, st, r, h, visits, tests, SceneCoverage{&S})) continue;
    const V3 pos = r.o + r.d * h.t;
    float clip[4];
    mul_point(F.world_to_clip, pos.x, pos.y, pos.z, 1.0f, clip);
    if (depth) depth[slot] = clip[2] / clip[3];
    if (!write_color) continue;
    float4 c;
    if (F.output == 1u) {
      c = make_float4(1.0f - h.u - h.v, h.u, h.v, 1.0f);
    } else {
      const Surface sf = reconstruct_hit(S, h.tri, h.instance, h.u, h.v, r.d);
      if (F.output == 3u) c = make_float4(sf.base_color.x, sf.base_color.y, sf.base_color.z, 1.0f);
      else if (F.output == 4u) c = make_float4(sf.roughness, sf.metallic, 0.0f, 1.0f);
      else c = make_float4(sf.shading_normal.x, sf.shading_normal.y, sf.shading_normal.z, 1.0f);
    }
    q.color[slot] = c;
  }
}

// Trace / ShadowTransmittance over caller rays (8 floats in, 5 words or 1 out).
__global__ void __launch_bounds__(kTraceBlock) k_rays(SceneView S, const float4* rays, uint32_t* hits,
                                                      uint32_t count, uint32_t any_hit, uint2* spill, int spill_cap) {
  __shared__ uint2 ring[kRing * kTraceBlock];
  TraceStack st;
  st.init(ring, spill, spill_cap);
  strided_walks(
      S, st, count,
      [&](uint32_t i, auto& w) {
        const float4 o4 = rays[2 * i], d4 = rays[2 * i + 1];
        w.start(Ray{xyz(o4), o4.w, xyz(d4), d4.w}, any_hit != 0);
      },
      [&](uint32_t i, const auto& w) {
        if (any_hit == 2u) {
          hits[i] = __float_as_uint(w.transmittance());
          return;
        }
        if (any_hit) {
          hits[i] = w.transmittance() > 0.0f ? 0u : 1u;
          return;
        }
        uint32_t* o = hits + 5ull * i;
        if (w.hit()) {
          const HitRecord h = w.record(S.tris);
          o[0] = h.instance;
          o[1] = h.primitive;
          o[2] = __float_as_uint(h.u);
          o[3] = __float_as_uint(h.v);
          o[4] = __float_as_uint(h.t);
        } else {
          o[0] = kMiss;
          o[1] = kMiss;
          o[2] = 0u;
          o[3] = 0u;
          o[4] = 0u;
        }
      },
      SceneCoverage{&S});
}

__global__ void k_zero(uint32_t* p, uint32_t n) {
  for (uint32_t i = gtid(); i < n; i += gstride()) p[i] = 0u;
}

__global__ void k_fill4(float4* p, uint32_t n, float4 v) {
  for (uint32_t i = gtid(); i < n; i += gstride()) p[i] = v;
}

__global__ void k_fill1(float* p, uint32_t n, float v) {
  for (uint32_t i = gtid(); i < n; i += gstride()) p[i] = v;
}

// Workgroups that fill every CU at the kernel's occupancy.
template <typename K>
uint32_t persistent_blocks(K kernel, int block, uint32_t cus) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1)
    per_cu = 2;
  return cus * static_cast<uint32_t>(per_cu);
}

uint32_t small_grid(uint32_t n) { return (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024; }

}  // namespace

GridShape grid_shape(uint32_t cus) {
  GridShape g;
  g.camera = persistent_blocks(k_camera<false, false>, kTraceBlock, cus);
  g.camera_dynamic = persistent_blocks(k_camera<true, false>, kTraceBlock, cus);
  // sized for the variants of scenes without analytic lights or covered surfaces
  g.trace = persistent_blocks(k_trace<false, false>, kTraceBlock, cus);
  g.trace_dynamic = persistent_blocks(k_trace<true, false>, kTraceBlock, cus);
  g.tail = persistent_blocks(k_tail<false, false, false>, kTraceBlock, cus);
  g.tail_cache = persistent_blocks(k_tail<true, false, false>, kTraceBlock, cus);
  g.camera_pass = persistent_blocks(k_camera_pass, kTraceBlock, cus);
  g.paths_variant[0] = persistent_blocks(k_paths_views<false, false>, kTraceBlock, cus);
  g.paths_variant[1] = persistent_blocks(k_paths_views<true, false>, kTraceBlock, cus);
  g.paths_variant[2] = persistent_blocks(k_paths_views<true, true>, kTraceBlock, cus);
  g.paths = std::max({g.paths_variant[0], g.paths_variant[1], g.paths_variant[2]});
  g.rays = persistent_blocks(k_rays, kTraceBlock, cus);
  g.shade = persistent_blocks(k_shade<false, false>, kShadeBlock, cus);
  g.shade_parallel = persistent_blocks(k_shade<true, false, true>, kShadeBlock, cus);
  g.resolve = cus * 4;
  g.max_trace_lanes =
      std::max({g.camera, g.camera_dynamic, g.trace, g.trace_dynamic, g.tail, g.tail_cache, g.camera_pass, g.rays,
                g.shade_parallel, g.paths}) * kTraceBlock;
  return g;
}

void launch_zero(uint32_t* p, uint32_t n, hipStream_t s) { k_zero<<<small_grid(n), 256, 0, s>>>(p, n); }
void launch_fill4(float4* p, uint32_t n, float4 v, hipStream_t s) { k_fill4<<<small_grid(n), 256, 0, s>>>(p, n, v); }
void launch_fill1(float* p, uint32_t n, float v, hipStream_t s) { k_fill1<<<small_grid(n), 256, 0, s>>>(p, n, v); }
void launch_camera(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c) {
  const bool cov = S.covered_instances != 0;
  const uint32_t t = L.camera_tile_shift;
  if (L.dynamic_camera && cov)
    k_camera<true, true><<<L.grid.camera_dynamic, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats,
                                                                             t);
  else if (L.dynamic_camera)
    k_camera<true, false><<<L.grid.camera_dynamic, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap,
                                                                              L.stats, t);
  else if (cov)
    k_camera<false, true><<<L.grid.camera, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats, t);
  else
    k_camera<false, false><<<L.grid.camera, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats, t);
}
void launch_shade(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                  uint32_t bounce, bool link) {
  const uint32_t l = link ? 1u : 0u;
  if (S.gltf_materials != 0) {
    // glTF PBR scenes: the variants with its closure
    const bool cov = S.covered_instances != 0, par = parallel_walks(S, F);
    const uint32_t grid = par ? L.grid.shade_parallel : L.grid.shade;
    if (par && cov)
      k_shade<true, true, true, true><<<grid, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce, l, L.spill, L.spill_cap);
    else if (par)
      k_shade<true, false, true, true><<<grid, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce, l, L.spill, L.spill_cap);
    else if (cov)
      k_shade<true, true, false, true><<<grid, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce, l, nullptr, 0);
    else
      k_shade<true, false, false, true><<<grid, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce, l, nullptr, 0);
  } else if (parallel_walks(S, F)) {
    if (S.covered_instances != 0)
      k_shade<true, true, true><<<L.grid.shade_parallel, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce, l, L.spill,
                                                                                     L.spill_cap);
    else
      k_shade<true, false, true><<<L.grid.shade_parallel, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce, l, L.spill,
                                                                                      L.spill_cap);
  } else if (S.covered_instances != 0) {
    k_shade<true, true><<<L.grid.shade, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce, l, nullptr, 0);
  } else if (scene_has_extras(S)) {
    k_shade<true, false><<<L.grid.shade, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce, l, nullptr, 0);
  } else {
    k_shade<false, false><<<L.grid.shade, kShadeBlock, 0, L.stream>>>(S, F, q, c, bounce, l, nullptr, 0);
  }
}
void launch_trace(const Launch& L, const SceneView& S, const Queues& q, uint32_t* c, uint32_t bounce,
                  uint32_t kinds) {
  const bool cov = S.covered_instances != 0;
  if (L.dynamic_trace && cov)
    k_trace<true, true><<<L.grid.trace_dynamic, kTraceBlock, 0, L.stream>>>(S, q, c, L.spill, L.spill_cap, bounce,
                                                                          kinds, L.stats);
  else if (L.dynamic_trace)
    k_trace<true, false><<<L.grid.trace_dynamic, kTraceBlock, 0, L.stream>>>(S, q, c, L.spill, L.spill_cap, bounce,
                                                                           kinds, L.stats);
  else if (cov)
    k_trace<false, true><<<L.grid.trace, kTraceBlock, 0, L.stream>>>(S, q, c, L.spill, L.spill_cap, bounce, kinds,
                                                                    L.stats);
  else
    k_trace<false, false><<<L.grid.trace, kTraceBlock, 0, L.stream>>>(S, q, c, L.spill, L.spill_cap, bounce, kinds,
                                                                     L.stats);
}
void launch_escape(const Launch& L, const SceneView& S, const Queues& q, uint32_t* c, uint32_t bounce, bool weigh) {
  k_escape<<<L.grid.resolve, 256, 0, L.stream>>>(S, q, c, bounce, weigh ? 1u : 0u);
}
void launch_tail(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c,
                 uint32_t bounce) {
  const bool analytic = scene_has_extras(S), cov = S.covered_instances != 0;
  const uint32_t grid = F.cache_depth != 0 ? L.grid.tail_cache : L.grid.tail;
  if (S.gltf_materials != 0) {
    // glTF PBR scenes: the variants with its closure
    const bool cache = F.cache_depth != 0, par = parallel_walks(S, F);
    if (cache) {
      if (par && cov)
        k_tail<true, true, true, true, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
      else if (par)
        k_tail<true, true, false, true, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
      else if (cov)
        k_tail<true, true, true, false, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
      else
        k_tail<true, true, false, false, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
    } else {
      if (par && cov)
        k_tail<false, true, true, true, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
      else if (par)
        k_tail<false, true, false, true, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
      else if (cov)
        k_tail<false, true, true, false, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
      else
        k_tail<false, true, false, false, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
    }
  } else if (parallel_walks(S, F)) {
    if (F.cache_depth != 0 && cov)
      k_tail<true, true, true, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce,
                                                                         L.stats);
    else if (F.cache_depth != 0)
      k_tail<true, true, false, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce,
                                                                          L.stats);
    else if (cov)
      k_tail<false, true, true, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce,
                                                                          L.stats);
    else
      k_tail<false, true, false, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce,
                                                                           L.stats);
  } else if (F.cache_depth != 0 && cov)
    k_tail<true, true, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
  else if (F.cache_depth != 0 && analytic)
    k_tail<true, true, false><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
  else if (F.cache_depth != 0)
    k_tail<true, false, false><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
  else if (cov)
    k_tail<false, true, true><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
  else if (analytic)
    k_tail<false, true, false><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
  else
    k_tail<false, false, false><<<grid, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, bounce, L.stats);
}
bool paths_eligible(const SceneView& S, const FrameConst& F) {
  return F.light_sampling <= 1u && F.cache_depth == 0 && F.restir_surf == nullptr && S.gltf_materials == 0 &&
         !parallel_walks(S, F);
}
void launch_paths(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c) {
  const bool analytic = scene_has_extras(S), cov = S.covered_instances != 0;
  const uint32_t* g = L.grid.paths_variant;
  if (cov)
    k_paths<true, true><<<g[2], kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats);
  else if (analytic)
    k_paths<true, false><<<g[1], kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats);
  else
    k_paths<false, false><<<g[0], kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats);
}
void launch_paths_views(const Launch& L, const PathView* views, uint32_t view_count, bool analytic, bool cov,
                        uint32_t* next_chunk, uint32_t total_chunks) {
  // every variant with the covered variant's grid, the smallest: more
  // resident waves hold chunks of more scenes at once and thrash the L2s
  // (measured: 3 workgroups per CU instead of 2 made the 64-view batch 22 %
  // slower)
  const uint32_t g = L.grid.paths_variant[2];
  if (cov)
    k_paths_views<true, true><<<g, kTraceBlock, 0, L.stream>>>(views, view_count, next_chunk, total_chunks,
                                                               L.spill, L.spill_cap);
  else if (analytic)
    k_paths_views<true, false><<<g, kTraceBlock, 0, L.stream>>>(views, view_count, next_chunk, total_chunks,
                                                                L.spill, L.spill_cap);
  else
    k_paths_views<false, false><<<g, kTraceBlock, 0, L.stream>>>(views, view_count, next_chunk, total_chunks,
                                                                 L.spill, L.spill_cap);
}
uint32_t path_count(const FrameConst& F) { return window_paths(F.rect[2], F.rect[3]); }
uint32_t path_chunks(const FrameConst& F) { return (path_count(F) + kPathChunk - 1) / kPathChunk; }
void launch_resolve(const Launch& L, const FrameConst& F, const Queues& q, uint32_t add_sample) {
  k_resolve<<<L.grid.resolve, 256, 0, L.stream>>>(F, q, add_sample);
}
void launch_camera_pass(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, float* depth,
                        uint32_t write_color) {
  k_camera_pass<<<L.grid.camera_pass, kTraceBlock, 0, L.stream>>>(S, F, q, depth, write_color, L.spill,
                                                                   L.spill_cap);
}
void launch_restir_initial(const Launch& L, const SceneView& S, const FrameConst& F, const Queues& q, uint32_t* c) {
  if (S.gltf_materials != 0)
    k_restir_initial<true, true><<<L.grid.tail, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats);
  else if (scene_has_extras(S))
    k_restir_initial<true><<<L.grid.tail, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats);
  else
    k_restir_initial<false><<<L.grid.tail, kTraceBlock, 0, L.stream>>>(S, F, q, c, L.spill, L.spill_cap, L.stats);
}
void launch_restir_clear(const Launch& L, float4* restir, float4* surf, uint32_t pixels) {
  k_restir_clear<<<L.grid.resolve, 256, 0, L.stream>>>(restir, surf, pixels);
}
void launch_restir_temporal(const Launch& L, const SceneView& S, const FrameConst& F, const RestirReuseArgs& A,
                            uint32_t* c) {
  if (S.gltf_materials != 0)
    k_restir_temporal<true, true><<<L.grid.tail, kTraceBlock, 0, L.stream>>>(S, F, A, c, L.spill, L.spill_cap,
                                                                            L.stats);
  else if (scene_has_extras(S))
    k_restir_temporal<true><<<L.grid.tail, kTraceBlock, 0, L.stream>>>(S, F, A, c, L.spill, L.spill_cap, L.stats);
  else
    k_restir_temporal<false><<<L.grid.tail, kTraceBlock, 0, L.stream>>>(S, F, A, c, L.spill, L.spill_cap, L.stats);
}
void launch_restir_spatial(const Launch& L, const SceneView& S, const FrameConst& F, const RestirReuseArgs& A,
                           uint32_t* c) {
  if (S.gltf_materials != 0)
    k_restir_spatial<true, true><<<L.grid.tail, kTraceBlock, 0, L.stream>>>(S, F, A, c, L.spill, L.spill_cap,
                                                                           L.stats);
  else if (scene_has_extras(S))
    k_restir_spatial<true><<<L.grid.tail, kTraceBlock, 0, L.stream>>>(S, F, A, c, L.spill, L.spill_cap, L.stats);
  else
    k_restir_spatial<false><<<L.grid.tail, kTraceBlock, 0, L.stream>>>(S, F, A, c, L.spill, L.spill_cap, L.stats);
}
void launch_cache_samples(const Launch& L, const FrameConst& F, const Queues& q, uint32_t* c, uint32_t capacity) {
  k_cache_samples<<<L.grid.resolve, 256, 0, L.stream>>>(F, q, c, capacity);
}
void launch_rays(const Launch& L, const SceneView& S, const float* rays, uint32_t* hits, uint32_t count,
                 uint32_t any_hit) {
  k_rays<<<L.grid.rays, kTraceBlock, 0, L.stream>>>(S, reinterpret_cast<const float4*>(rays), hits, count, any_hit,
                                                     L.spill, L.spill_cap);
}

}  // namespace lotus
