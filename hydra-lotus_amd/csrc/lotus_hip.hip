// Host side of liblotus_hip: the C ABI of include/lotus_hip.h. It plays the
// part of the reference's VulkanOffscreenRenderer for Integrator::Wavefront
// (backend/vulkan/vulkan_backend.cpp:1092-1399, 2399-2470, 2543-2619,
// 3745-3856) and of GpuScene::Apply (backend/vulkan/gpu_scene.cpp:1606-2450):
// scene upload and GPU BVH build, accumulation and restart rules, the
// wavefront schedule, per-kernel hipEvent timings and evidence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lotus_hip.h"
#include "products.hpp"
#include "temporal.hpp"

using namespace lotus;

namespace {

void set_detail(char* detail, size_t size, const std::string& s) {
  if (!detail || size == 0) return;
  std::strncpy(detail, s.c_str(), size - 1);
  detail[size - 1] = 0;
}

template <typename T>
hipError_t dev_alloc(T*& p, size_t count) {
  p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T));
}
template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// vulkan_internal.hpp:326-336.
void vulkan_world_to_clip(const float* gl, float* vk) {
  std::memcpy(vk, gl, 64);
  for (int c = 0; c < 4; ++c) {
    const float y = gl[c * 4 + 1], z = gl[c * 4 + 2], w = gl[c * 4 + 3];
    vk[c * 4 + 1] = -y;
    vk[c * 4 + 2] = 0.5f * (z + w);
  }
}

// vulkan_backend.cpp:436-478: Gauss-Jordan with partial pivoting in double.
bool invert(const float* m, float* inverse) {
  double rows[4][8] = {};
  for (int r = 0; r < 4; ++r) {
    for (int c = 0; c < 4; ++c) {
      const float v = m[c * 4 + r];
      if (!std::isfinite(v)) return false;
      rows[r][c] = v;
    }
    rows[r][r + 4] = 1.0;
  }
  for (int c = 0; c < 4; ++c) {
    int pivot = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(rows[r][c]) > std::fabs(rows[pivot][c])) pivot = r;
    if (rows[pivot][c] == 0.0) return false;
    for (int i = 0; i < 8; ++i) std::swap(rows[c][i], rows[pivot][i]);
    const double scale = rows[c][c];
    for (double& v : rows[c]) v /= scale;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = rows[r][c];
      for (int i = 0; i < 8; ++i) rows[r][i] -= f * rows[c][i];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      const float v = static_cast<float>(rows[r][c + 4]);
      if (!std::isfinite(v)) return false;
      inverse[c * 4 + r] = v;
    }
  return true;
}

// The rounds after a sample whose ray queues held `counts` paths per bounce
// (vulkan_backend.cpp:369-385): every bounce before the first whose queue
// held fewer than 1/divisor of the camera rays, and at least the camera
// rays' bounce, runs as a wavefront round; the megakernel tail follows the
// rest. A round's traversal launch lasts as long as its slowest ray, however
// few rays it holds (on the heightfield, grazing shadow rays walk hundreds of
// nodes), while the tail pays each path's slowest walks only once. Divisor 0:
// every bounce up to the last that traced a ray, plus one.
uint32_t wavefront_rounds(const std::vector<uint32_t>& counts, uint32_t divisor) {
  uint32_t rounds = 1;
  if (divisor == 0) {
    for (uint32_t b = 1; b < counts.size(); ++b)
      if (counts[b] != 0) rounds = b + 1;
    return rounds + 1;
  }
  const uint32_t camera = counts.empty() ? 0u : counts.front();
  const uint32_t threshold = std::max(1u, static_cast<uint32_t>((uint64_t{camera} + divisor - 1) / divisor));
  while (rounds < counts.size() && counts[rounds] >= threshold) ++rounds;
  return rounds;
}

// Rounds of a sample with no previous one to go by, and the most a sample
// records (the tail follows longer paths; the event pool holds their timers).
constexpr uint32_t kFirstRounds = 32;
constexpr uint32_t kMaxRounds = 64;

struct AccumulationKey {  // vulkan_backend.cpp:419-434
  std::array<float, 16> world_to_clip{};
  uint32_t width = 0, height = 0;
  std::array<float, 4> display{};
  std::array<uint32_t, 4> data{};
  uint32_t light_sampling = 0, sample_index = 0, max_bounces = 0;
  uint64_t scene_generation = 0;
  uint32_t cache_frames = 0;  // radiance cache frames accumulate apart
  uint32_t restir_candidates = 0, restir_proposal = 0;
  bool operator==(const AccumulationKey& o) const {
    return world_to_clip == o.world_to_clip && width == o.width && height == o.height && display == o.display &&
           data == o.data && light_sampling == o.light_sampling && sample_index == o.sample_index &&
           max_bounces == o.max_bounces && scene_generation == o.scene_generation && cache_frames == o.cache_frames &&
           restir_candidates == o.restir_candidates && restir_proposal == o.restir_proposal;
  }
};

enum Timer {
  kTCamera, kTIntersect, kTShade, kTShadow, kTTrace, kTTail, kTResolve, kTCacheInfer, kTCacheTrain, kTRestir,
  kTTemporal, kTimers
};

// One temporal frame's device work and buffers (lotus_hip_render_temporal).
struct TemporalFrameState {
  TemporalConst T{};
  TemporalInstances I{};
  float view[16];
  bool reuse = false, spatial = false, accumulate = false;
  bool fresh = false;              // a fresh estimator batch (accumulation or reservoir reuse)
  bool settling = false;           // SettleTemporalRadiance instead of the accumulation and filter
  const float4* prior = nullptr;   // the last changing frame's colour
  uint32_t max_samples = 0;
  float* batch_host = nullptr;     // the batch product (resolved colour before the SVGF passes)
  RestirReuseArgs A{};
  TemporalPixel* surface = nullptr;
  const TemporalPixel* previous_surface = nullptr;
  TemporalValidityPixel* validity = nullptr;
  TemporalRadiancePixel* radiance = nullptr;
  const TemporalRadiancePixel* old_radiance = nullptr;
  float4* filter = nullptr;
};

// gpu_scene.cpp:144-179 (EmissionEstimate): the mean emission of a
// UsdPreviewSurface record, its lookup's max(scale, 0) + max(bias, 0) (or
// fallback) when its emission is textured; 0 unless finite and positive.
double emission_estimate(const uint32_t* m, const uint32_t* bindings) {
  double e[3];
  if ((m[2] & (1u << 5)) == 0) {
    for (int c = 0; c < 3; ++c) {
      float v;
      std::memcpy(&v, m + 16 + c, 4);
      e[c] = v;
    }
  } else {
    const uint32_t* b = bindings + 24ull * (m[1] + static_cast<uint32_t>(__builtin_popcount(m[2] & 31u)));
    float f[24];
    std::memcpy(f, b, sizeof f);
    for (int c = 0; c < 3; ++c)
      e[c] = b[3] == 2u ? std::max(0.0f, f[12 + c]) : double(std::max(0.0f, f[4 + c])) + std::max(0.0f, f[8 + c]);
  }
  // glTF PBR's emissive strength is a constant factor on the estimate
  double factor = 1.0;
  if ((m[0] & 0xFFu) == 1u) {
    float strength;
    std::memcpy(&strength, m + 20, 4);
    factor = strength;
  }
  const double mean = factor * (e[0] + e[1] + e[2]) / 3.0;
  return std::isfinite(mean) && mean > 0.0 ? mean : 0.0;
}

// The glTF closure's directional albedo tables (generated by
// hydra-lotus_amd/tables/make_albedo_tables.cpp in albedo.slang's layout).
const float kAlbedoTables[] = {
#include "albedo_tables.inc"
};

// The light table and the analytic lights as the device holds them.
struct LightTableHost {
  std::vector<uint32_t> offsets;  // per instance: its first entry or kNoLights
  std::vector<LightRecord> lights;
  std::vector<AnalyticRecord> analytic;
  std::vector<LightNode> nodes;
  std::vector<uint32_t> distant;
  uint32_t parallel_count = 0, parallel = 0;  // parallel lights; the first one's index + 1
  // The environment image (width 0: none): texels, its sampling
  // distribution (gpu_scene.cpp:631-704), its frame's axes, scale, and its
  // entry's probability.
  uint32_t env_width = 0, env_height = 0, env_columns = 0, env_rows = 0;
  std::vector<float4> env_texels;
  std::vector<float> env_dist;
  float env_frame[9] = {}, env_scale[3] = {}, env_selection = 0.0f;
  bool operator==(const LightTableHost& o) const {
    auto same = [](const auto& a, const auto& b) {
      return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(a[0])) == 0);
    };
    return same(offsets, o.offsets) && same(lights, o.lights) && same(analytic, o.analytic) && same(nodes, o.nodes) &&
           same(distant, o.distant) && env_width == o.env_width && env_height == o.env_height &&
           same(env_texels, o.env_texels) && same(env_dist, o.env_dist) &&
           std::memcmp(env_frame, o.env_frame, sizeof env_frame) == 0 &&
           std::memcmp(env_scale, o.env_scale, sizeof env_scale) == 0 && env_selection == o.env_selection;
  }
};

}  // namespace

struct lotus_hip_renderer {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t cus = 0;
  std::string failure;

  SceneBuffers scene;
  SceneView view{};
  bool has_scene = false;
  uint64_t scene_generation = 0;  // any change: restarts the accumulation
  uint64_t scene_epoch = 0;       // a change temporal history cannot carry over
  // The caller's per-instance temporal identities copied to the device:
  // previous_from_current (16 floats), instance and material ids, flags.
  unsigned char* temporal_ids = nullptr;
  size_t temporal_ids_bytes = 0;
  float* albedo_tables = nullptr;  // the glTF closure's albedo tables
  uint64_t emissive_triangles = 0;  // the light table's triangle entries
  LightTableHost light_table;     // as uploaded
  // The applied plan, to find what the next one changes.
  struct PlanRecord {
    std::vector<float> positions, transforms;
    std::vector<uint32_t> triangles, gvo, gvc, gto, gtc, igeo, imat, materials;
    std::vector<lotus_hip_light> lights;
    // texture lookups and per-corner data: texture sizes and formats, texel
    // bytes, bindings, each instance's first texture-coordinate, normal and
    // colour corner, and the coordinates, normals and colours
    std::vector<uint32_t> texture_shape, bindings, instance_texcoords, instance_normals, instance_colors;
    std::vector<uint8_t> texel_bytes;
    std::vector<float> texcoords, normals, colors;
    bool has_env = false;
    uint32_t env_width = 0, env_height = 0;
    std::vector<float> env_texels;
    float env_scale[3] = {0, 0, 0}, env_frame[16] = {};
    float environment[3] = {0, 0, 0};
    uint32_t leaf_size = 0;
    uint32_t builder = 0;
  } plan;
  // Lifetime scene counts (lotus_hip_scene_evidence).
  uint64_t bvh_builds = 0, bvh_refits = 0, material_writes = 0, light_writes = 0, uploaded_bytes = 0;

  uint32_t width = 0, height = 0;
  Queues q{};
  uint32_t* counters = nullptr;
  unsigned long long* stats = nullptr;  // 8: node visits / triangle tests per traversal kernel
  float* depth = nullptr;
  uint2* spill = nullptr;
  size_t spill_entries = 0;  // per lane
  GridShape grid{};
  lotus_hip_tuning tuning{0u, 0u, 16u, 0u, 0u, 0u};

  AccumulationKey key;
  bool key_valid = false;
  uint32_t accumulated = 0;
  uint32_t rounds_hint = 0;

  std::vector<hipEvent_t> events;

  // Temporal frames: ping-pong history buffers (index cur is this frame's)
  // and what the next frame compares against.
  struct Temporal {
    uint32_t width = 0, height = 0;
    TemporalPixel* surf[2] = {nullptr, nullptr};
    TemporalValidityPixel* validity = nullptr;
    TemporalRadiancePixel* rad[2] = {nullptr, nullptr};
    float4* filter = nullptr;
    float4* rsurf[2] = {nullptr, nullptr};
    float4* rhist[2] = {nullptr, nullptr};
    float4* rspin = nullptr;
    uint4* rsources = nullptr;
    float4* prior = nullptr;    // the last changing frame's colour, for settling
    uint32_t sample_offset = 0; // the fresh batches' samples since history was last dropped
    uint32_t cur = 0;
    uint32_t rcur = 0;          // the radiance history's write index (settling frames keep it)
    bool prior_valid = false;   // the last frame rendered with settle
    bool previous = false;      // the last successful temporal frame's state below
    uint64_t frame_index = 0, history_key = 0, scene_generation = 0, scene_epoch = 0;
    float view[16] = {}, world_to_clip[16] = {}, display[4] = {};
    float eye[3] = {};  // the camera's world-space position
    uint32_t data[4] = {};
    uint32_t light_sampling = 0, max_bounces = 0, candidates = 0, proposal = 0;
    bool reservoirs = false, radiance = false;
    // the windows as the caller gave them (TemporalHistory compares these)
    float raw_display[4] = {};
    int32_t raw_data[4] = {};
    // the last fresh batch's transport settings (TemporalTransportKey,
    // vulkan_backend.cpp:295-310); transport_valid: the last frame had one
    std::array<uint32_t, 14> transport{};
    bool transport_valid = false;
  } temporal;
  // A batched frame's sample whose counters are read after the batch.
  struct Deferred {
    bool pending = false;
    bool add = false;
    uint32_t stride = 0, rounds = 0;
    uint64_t launches = 0;  // its traversal launches
  } deferred;
  hipEvent_t join = nullptr;       // this renderer's branch of a batch ends
  // The executable graph of the last batch this renderer led (its first
  // view), for the views in `batch_views`: a later batch of the same views
  // updates it in place (hipGraphExecUpdate) instead of instantiating anew.
  hipGraphExec_t batch_exec = nullptr;
  std::vector<const lotus_hip_renderer*> batch_views;
  bool graph_batch = false;  // the batch this renderer's sample is enqueued in is captured
  // A captured sample's frame variables: their device words (in the batch
  // leader's array), this frame's values, and the sample's deferred record
  // and parameters, which a replay of the same graph reuses.
  uint32_t* graph_vars = nullptr;
  uint32_t graph_host_vars[4] = {0, 0, 0, 0};
  Deferred graph_deferred;
  std::vector<uint8_t> graph_signature;
  uint32_t* batch_vars = nullptr;  // the batch leader's frame variables, 4 words per view
  uint32_t batch_vars_count = 0;
  // A fused batch's view (render_impl stops before the paths and records
  // them here), and, held by the batch leader, the device view table, its
  // work-queue counters (one per kernel variant) and a traversal spill area
  // for the fused launches.
  bool fuse = false;
  bool fused_pending = false;
  PathView fused_view{};
  uint32_t fused_add = 0;
  PathView* fused_table = nullptr;
  uint32_t fused_table_count = 0;
  uint32_t* fused_queue = nullptr;  // kXcds words per kernel variant
  uint2* fused_spill = nullptr;
  size_t fused_spill_entries = 0;
  uint64_t batch_instantiations = 0, batch_updates = 0, batch_replays = 0;
  float4* svgf_filter = nullptr;   // lotus_hip_svgf's scratch
  uint32_t svgf_pixels = 0;
  float4* static_scratch = nullptr;  // this device's copy of `scratch` (lotus_hip_svgf_static)
  std::string static_scratch_error;

  // Radiance cache and its frame buffers.
  CacheState cache;
  lotus_hip_cache_config cache_config{};
  float scene_lo[3] = {0, 0, 0}, scene_scale[3] = {1, 1, 1};
  bool bounds_valid = false;
  uint32_t cache_pixels = 0;       // target size the frame buffers were made for
  uint32_t train_stride = 1, train_capacity = 0, sample_capacity = 0;
};

namespace {

void free_target(lotus_hip_renderer* r) {
  Queues& q = r->q;
  dev_free(q.thr);
  dev_free(q.rad);
  dev_free(q.nee);
  dev_free(q.esc);
  dev_free(q.esc2);
  dev_free(q.esc_o);
  dev_free(q.esc_d);
  dev_free(q.ray_o);
  dev_free(q.ray_d);
  dev_free(q.hit_o);
  dev_free(q.hit_d);
  dev_free(q.hit_h);
  dev_free(q.sh_o);
  dev_free(q.sh_d);
  dev_free(q.sh_c);
  dev_free(q.sh_f);
  dev_free(q.sh_t);
  dev_free(q.accum);
  dev_free(q.color);
  dev_free(q.cq);
  dev_free(q.cq_thr);
  dev_free(q.crec);
  dev_free(q.crec_n);
  dev_free(q.samples);
  dev_free(q.restir);
  dev_free(r->depth);
  r->width = r->height = 0;
  r->cache_pixels = 0;
}

void free_scene(lotus_hip_renderer* r) {
  SceneBuffers& s = r->scene;
  dev_free(s.nodes);
  dev_free(s.tris);
  dev_free(s.tri_of_global);
  dev_free(s.instance_base);
  dev_free(s.instance_material);
  dev_free(s.instance_transform);
  dev_free(s.materials);
  dev_free(s.light_offsets);
  dev_free(s.lights);
  dev_free(s.analytic);
  dev_free(s.light_nodes);
  dev_free(s.distant);
  dev_free(s.env_texels);
  dev_free(s.env_dist);
  dev_free(s.textures);
  dev_free(s.texels);
  dev_free(s.bindings);
  dev_free(s.texcoords);
  dev_free(s.instance_texcoords);
  dev_free(s.normals);
  dev_free(s.instance_normals);
  dev_free(s.instance_covered);
  s.covered_host.clear();
  s.covered_instances = 0;
  dev_free(s.colors);
  dev_free(s.instance_colors);
  dev_free(s.positions);
  dev_free(s.triangles);
  dev_free(s.geometry_vertex_offset);
  dev_free(s.geometry_triangle_offset);
  dev_free(s.instance_geometry);
  dev_free(s.exact);
  s = SceneBuffers{};
  r->plan = lotus_hip_renderer::PlanRecord{};
  r->light_table = LightTableHost{};
  r->has_scene = false;
}

// Persistent grids and the traversal stacks' spill area, sized for the BVH's depth.
// A nearest-first walk of a 4-wide tree keeps at most 3 entries per level.
hipError_t ensure_launch(lotus_hip_renderer* r) {
  const size_t stack = 3 * size_t(r->scene.depth) + 4;
  const size_t need = std::max<size_t>(4, stack > size_t(kRing) ? stack - kRing : 0);
  if (r->spill && r->spill_entries >= need) return hipSuccess;
  dev_free(r->spill);
  r->spill_entries = need;
  return dev_alloc(r->spill, need * size_t(r->grid.max_trace_lanes));
}

// The launch shapes under the renderer's tuning: the persistent traversal
// grids scaled down to the resident share (never up, so the spill area sized
// for the full grids holds every lane).
Launch launch_for(lotus_hip_renderer* r, hipStream_t stream) {
  GridShape g = r->grid;
  const uint32_t pct = r->tuning.resident_percent;
  if (pct != 0 && pct < 100) {
    for (uint32_t* n : {&g.camera, &g.camera_dynamic, &g.trace, &g.trace_dynamic, &g.tail, &g.tail_cache,
                        &g.paths_variant[0], &g.paths_variant[1], &g.paths_variant[2]})
      *n = std::max(1u, static_cast<uint32_t>((uint64_t{*n} * pct) / 100));
  }
  // camera tiles: 8 x 8 (0), 16 x 4 (1) or 4 x 16 (2)
  const uint32_t shift = r->tuning.camera_tile == 1 ? 4u : r->tuning.camera_tile == 2 ? 2u : 3u;
  return Launch{stream, g, r->spill, static_cast<int>(r->spill_entries), r->stats,
                r->tuning.camera_dynamic != 0, r->tuning.trace_dynamic != 0, shift};
}

hipError_t ensure_target(lotus_hip_renderer* r, uint32_t width, uint32_t height) {
  if (r->width == width && r->height == height && r->q.thr) return hipSuccess;
  free_target(r);
  const size_t P = size_t(width) * height;
  Queues& q = r->q;
  hipError_t e = hipSuccess;
  for (float4** p : {&q.thr, &q.rad, &q.nee, &q.esc, &q.esc2, &q.esc_o, &q.esc_d, &q.ray_o, &q.ray_d, &q.hit_o, &q.hit_d, &q.sh_o, &q.sh_d, &q.sh_c, &q.sh_f,
                     &q.sh_t, &q.accum, &q.color})
    if (e == hipSuccess) e = dev_alloc(*p, P);
  if (e == hipSuccess) e = dev_alloc(q.hit_h, P);
  if (e == hipSuccess) e = dev_alloc(r->depth, P);
  if (e != hipSuccess) {
    free_target(r);
    return e;
  }
  r->width = width;
  r->height = height;
  r->key_valid = false;
  return hipSuccess;
}

// Radiance cache frame buffers: queries for every pixel, training paths one
// pixel in train_stride, their recorded vertices and samples.
hipError_t ensure_cache_target(lotus_hip_renderer* r) {
  const uint32_t P = r->width * r->height;
  const uint32_t paths = std::max(1u, r->cache_config.train_paths);
  const uint32_t stride = std::max(1u, P / paths);
  if (r->q.cq && r->cache_pixels == P && r->train_stride == stride) return hipSuccess;
  Queues& q = r->q;
  dev_free(q.cq);
  dev_free(q.cq_thr);
  dev_free(q.crec);
  dev_free(q.crec_n);
  dev_free(q.samples);
  r->cache_pixels = 0;
  r->train_stride = stride;
  r->train_capacity = P / stride + 2;
  r->sample_capacity = r->train_capacity * kCacheRecordDepth;
  hipError_t e = dev_alloc(q.cq, 4ull * P);
  if (e == hipSuccess) e = dev_alloc(q.cq_thr, P);
  if (e == hipSuccess) e = dev_alloc(q.crec, 6ull * kCacheRecordDepth * r->train_capacity);
  if (e == hipSuccess) e = dev_alloc(q.crec_n, r->train_capacity);
  if (e == hipSuccess) e = dev_alloc(q.samples, size_t(kSampleFloats) * r->sample_capacity);
  if (e == hipSuccess) r->cache_pixels = P;
  return e;
}

// The scene's box (the root's child boxes) mapped to the cache's unit cube.
hipError_t ensure_bounds(lotus_hip_renderer* r) {
  if (r->bounds_valid) return hipSuccess;
  float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  if (r->scene.node_count != 0) {
    WideNode root;
    const hipError_t e = hipMemcpy(&root, r->scene.nodes, sizeof root, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    const float* los[3] = {root.lox, root.loy, root.loz};
    const float* his[3] = {root.hix, root.hiy, root.hiz};
    for (int a = 0; a < 3; ++a) {
      float l = INFINITY, h = -INFINITY;
      for (int c = 0; c < 4; ++c)
        if (los[a][c] <= his[a][c]) l = std::min(l, los[a][c]), h = std::max(h, his[a][c]);
      if (l <= h) lo[a] = l, hi[a] = h;
    }
  }
  for (int a = 0; a < 3; ++a) {
    const float ext = hi[a] - lo[a];
    r->scene_lo[a] = lo[a] - 1e-4f * std::max(ext, 1e-3f);
    r->scene_scale[a] = 1.0f / (std::max(ext, 1e-3f) * 1.0002f);
  }
  r->bounds_valid = true;
  return hipSuccess;
}

uint32_t mix_host(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

void free_temporal(lotus_hip_renderer* r) {
  auto& t = r->temporal;
  for (int k = 0; k < 2; ++k) {
    dev_free(t.surf[k]);
    dev_free(t.rad[k]);
    dev_free(t.rsurf[k]);
    dev_free(t.rhist[k]);
  }
  dev_free(t.validity);
  dev_free(t.filter);
  dev_free(t.rspin);
  dev_free(t.rsources);
  dev_free(t.prior);
  t.width = t.height = 0;
  t.previous = false;
  t.prior_valid = false;
}

hipError_t ensure_temporal(lotus_hip_renderer* r, uint32_t width, uint32_t height) {
  auto& t = r->temporal;
  if (t.width == width && t.height == height && t.surf[0]) return hipSuccess;
  free_temporal(r);
  const size_t P = size_t(width) * height;
  hipError_t e = hipSuccess;
  for (int k = 0; k < 2 && e == hipSuccess; ++k) {
    e = dev_alloc(t.surf[k], P);
    if (e == hipSuccess) e = dev_alloc(t.rad[k], P);
    if (e == hipSuccess) e = dev_alloc(t.rsurf[k], kReuseSurfaceWords * P);
    if (e == hipSuccess) e = dev_alloc(t.rhist[k], 4 * P);
  }
  if (e == hipSuccess) e = dev_alloc(t.validity, P);
  if (e == hipSuccess) e = dev_alloc(t.filter, 4 * P);
  if (e == hipSuccess) e = dev_alloc(t.rspin, 4 * P);
  if (e == hipSuccess) e = dev_alloc(t.rsources, 2 * P);
  if (e == hipSuccess) e = dev_alloc(t.prior, P);
  if (e != hipSuccess) {
    free_temporal(r);
    return e;
  }
  t.width = width;
  t.height = height;
  return hipSuccess;
}

std::string latch(lotus_hip_renderer* r, const std::string& what, hipError_t e) {
  r->failure = what + ": " + hipGetErrorString(e);
  return r->failure;
}

void store4(float* out, double x, double y, double z, double w) {
  out[0] = static_cast<float>(x);
  out[1] = static_cast<float>(y);
  out[2] = static_cast<float>(z);
  out[3] = static_cast<float>(w);
}

// gpu_scene.cpp:233-270 (SceneRadius): half the diagonal of the world box of
// the instances' object-space bounds' corners, in double.
double scene_radius(const lotus_hip_scene_plan* p) {
  const double inf = std::numeric_limits<double>::infinity();
  std::vector<std::array<double, 6>> object(p->geometry_count, {inf, inf, inf, -inf, -inf, -inf});
  for (uint32_t g = 0; g < p->geometry_count; ++g) {
    const float* pos = p->positions + 3ull * p->geometry_vertex_offset[g];
    for (uint32_t v = 0; v < p->geometry_vertex_count[g]; ++v)
      for (int a = 0; a < 3; ++a) {
        object[g][a] = std::min(object[g][a], double(pos[3 * v + a]));
        object[g][3 + a] = std::max(object[g][3 + a], double(pos[3 * v + a]));
      }
  }
  std::array<double, 6> world{inf, inf, inf, -inf, -inf, -inf};
  for (uint32_t i = 0; i < p->instance_count; ++i) {
    const auto& b = object[p->instance_geometry[i]];
    if (!(b[0] <= b[3])) continue;
    const float* m = p->instance_transform + 16ull * i;
    for (int corner = 0; corner < 8; ++corner) {
      const double q[3] = {b[(corner & 1) ? 3 : 0], b[(corner & 2) ? 4 : 1], b[(corner & 4) ? 5 : 2]};
      for (int r = 0; r < 3; ++r) {
        const double v = double(m[r]) * q[0] + double(m[4 + r]) * q[1] + double(m[8 + r]) * q[2] + double(m[12 + r]);
        world[r] = std::min(world[r], v);
        world[3 + r] = std::max(world[3 + r], v);
      }
    }
  }
  if (!(world[0] <= world[3])) return 0.0;
  const double dx = world[3] - world[0], dy = world[4] - world[1], dz = world[5] - world[2];
  const double radius = 0.5 * std::sqrt(dx * dx + dy * dy + dz * dz);
  return std::isfinite(radius) ? radius : 0.0;
}

// gpu_scene.cpp:271-369 (MakeAnalyticLight): the world-space record and the
// light-table weight, the estimated power over 2 pi (a distant light's over
// a disk of the scene's radius). False for a light that emits nothing.
bool make_analytic_light(const lotus_hip_light& L, double scene_r, AnalyticRecord& rec, double& weight) {
  constexpr double kPiD = 3.14159265358979323846;
  const float* m = L.world_from_light;
  const double x[3] = {m[0], m[1], m[2]}, y[3] = {m[4], m[5], m[6]}, z[3] = {m[8], m[9], m[10]},
               o[3] = {m[12], m[13], m[14]};
  auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  auto cross = [](const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
  };
  double yz[3];
  cross(y, z, yz);
  const double det = dot(x, yz);
  rec = AnalyticRecord{};
  double extent = 0.0, scale = 1.0;
  if (L.shape == LOTUS_HIP_LIGHT_SPHERE) {
    const double radius = double(L.radius) * std::cbrt(std::fabs(det));
    extent = 4.0 * kPiD * radius * radius;
    rec.shape = kSphereLight;
    store4(rec.position, o[0], o[1], o[2], radius);
    store4(rec.normal, 0.0, 0.0, 0.0, extent);
  } else if (L.shape == LOTUS_HIP_LIGHT_RECT || L.shape == LOTUS_HIP_LIGHT_DISK) {
    const bool rect = L.shape == LOTUS_HIP_LIGHT_RECT;
    const double su = rect ? 0.5 * double(L.width) : double(L.radius);
    const double sv = rect ? 0.5 * double(L.height) : double(L.radius);
    const double u[3] = {x[0] * su, x[1] * su, x[2] * su}, v[3] = {y[0] * sv, y[1] * sv, y[2] * sv};
    double n[3];
    cross(u, v, n);
    const double len = std::sqrt(dot(n, n));
    extent = (rect ? 4.0 : kPiD) * len;
    rec.shape = rect ? kRectLight : kDiskLight;
    store4(rec.position, o[0], o[1], o[2], 0.0);
    store4(rec.axis_u, u[0], u[1], u[2], 0.0);
    store4(rec.axis_v, v[0], v[1], v[2], 0.0);
    if (len > 0.0) {
      const double k = (det < 0.0 ? 1.0 : -1.0) / len;  // the -Z side, through the inverse transpose
      store4(rec.normal, n[0] * k, n[1] * k, n[2] * k, extent);
    }
  } else {  // distant
    const double len = std::sqrt(dot(z, z));
    if (!(len > 0.0)) return false;
    const double theta = std::min(double(L.angle) * kPiD / 360.0, kPiD);
    const double half = std::sin(0.5 * theta);
    const double one_minus_cos = 2.0 * half * half;
    const double sin2 = std::sin(theta) * std::sin(theta);
    store4(rec.axis_u, z[0] / len, z[1] / len, z[2] / len, one_minus_cos);
    rec.shape = kParallelLight;
    extent = 1.0;
    if (theta > 0.0) {
      extent = theta <= 0.5 * kPiD ? kPiD * sin2 : kPiD;
      if (L.normalize) scale = 1.0 / (theta <= 0.5 * kPiD ? kPiD * sin2 : (2.0 - sin2) * kPiD);
      if (static_cast<float>(one_minus_cos) >= std::numeric_limits<float>::min()) {
        rec.shape = kDistantLight;
      } else {  // too narrow a cone for float densities: one direction
        scale *= extent;
        extent = 1.0;
      }
    }
  }
  if (L.normalize && L.shape != LOTUS_HIP_LIGHT_DISTANT) scale = 1.0 / extent;
  double radiance[3];
  for (int c = 0; c < 3; ++c) radiance[c] = double(L.radiance[c]) * scale;
  const double mean = (radiance[0] + radiance[1] + radiance[2]) / 3.0;
  const double fmax = std::numeric_limits<float>::max();
  if (!(extent > 0.0) || !std::isfinite(extent) || !(mean > 0.0) || !(radiance[0] <= fmax) ||
      !(radiance[1] <= fmax) || !(radiance[2] <= fmax))
    return false;
  store4(rec.radiance, radiance[0], radiance[1], radiance[2], 0.0);
  weight = mean * extent / 2.0;
  if (L.shape == LOTUS_HIP_LIGHT_DISTANT) weight *= scene_r * scene_r;
  if (!std::isfinite(weight)) weight = std::numeric_limits<double>::max();
  return true;
}

// gpu_scene.cpp:379-442: a bounded light's box, widened by 1e-4 of its extent
// plus 1e-5 of its coordinates' magnitude and rounded outwards.
struct LightBox {
  float low[3], high[3];
  double centre[3];
  uint32_t index;
};

LightBox bound_light(const AnalyticRecord& l, uint32_t index) {
  LightBox b;
  b.index = index;
  for (int a = 0; a < 3; ++a) {
    const double c = l.position[a], u = l.axis_u[a], v = l.axis_v[a];
    const double extent = l.shape == kSphereLight ? double(l.position[3])
                          : l.shape == kRectLight ? std::fabs(u) + std::fabs(v)
                                                  : std::sqrt(u * u + v * v);
    const double margin = 1e-4 * extent + 1e-5 * (std::fabs(c) + extent);
    b.centre[a] = c;
    b.low[a] = std::nextafter(static_cast<float>(c - extent - margin), -std::numeric_limits<float>::infinity());
    b.high[a] = std::nextafter(static_cast<float>(c + extent + margin), std::numeric_limits<float>::infinity());
  }
  return b;
}

void store_bounds(float* low, float* high, const LightBox* first, const LightBox* last, uint32_t child) {
  for (int a = 0; a < 3; ++a) {
    low[a] = std::numeric_limits<float>::max();
    high[a] = -std::numeric_limits<float>::max();
    for (const LightBox* b = first; b != last; ++b) {
      low[a] = std::min(low[a], b->low[a]);
      high[a] = std::max(high[a], b->high[a]);
    }
  }
  std::memcpy(&low[3], &child, 4);
}

// A median split along the axis the centres spread most over, ties by index
// (gpu_scene.cpp:405-441): the hierarchy depends only on the lights.
uint32_t build_light_node(LightBox* first, LightBox* last, std::vector<LightNode>& nodes) {
  if (last - first == 1) return kLightLeaf | first->index;
  const uint32_t node = static_cast<uint32_t>(nodes.size());
  nodes.emplace_back();
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = first->centre[a];
  for (const LightBox* b = first; b != last; ++b)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b->centre[a]);
      hi[a] = std::max(hi[a], b->centre[a]);
    }
  int axis = 0;
  for (int a = 1; a < 3; ++a)
    if (hi[a] - lo[a] > hi[axis] - lo[axis]) axis = a;
  LightBox* middle = first + (last - first) / 2;
  std::nth_element(first, middle, last, [axis](const LightBox& a, const LightBox& b) {
    return a.centre[axis] != b.centre[axis] ? a.centre[axis] < b.centre[axis] : a.index < b.index;
  });
  const uint32_t left = build_light_node(first, middle, nodes);
  const uint32_t right = build_light_node(middle, last, nodes);
  LightNode& n = nodes[node];
  store_bounds(n.low0, n.high0, first, middle, left);
  store_bounds(n.low1, n.high1, middle, last, right);
  return node;
}

// render_world.cpp:314-345 (EnvironmentFrame): the frame's X, Y, Z columns
// orthonormalized in that order; false when they are not independent.
bool environment_frame(const float* m, double axes[3][3]) {
  for (int k = 0; k < 16; ++k)
    if (!std::isfinite(m[k])) return false;
  for (int column = 0; column < 3; ++column) {
    double axis[3] = {m[4 * column], m[4 * column + 1], m[4 * column + 2]};
    const double length = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    for (int previous = 0; previous < column; ++previous) {
      const double along = axis[0] * axes[previous][0] + axis[1] * axes[previous][1] + axis[2] * axes[previous][2];
      for (int r = 0; r < 3; ++r) axis[r] -= along * axes[previous][r];
    }
    const double remainder = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    if (!std::isfinite(length) || !(remainder > 1e-6 * length)) return false;
    for (int r = 0; r < 3; ++r) axes[column][r] = axis[r] / remainder;
  }
  return true;
}

constexpr uint32_t kMaxEnvironmentColumns = 2048, kMaxEnvironmentRows = 1024;  // gpu_scene.hpp:211-212

// gpu_scene.cpp:631-704 (BuildEnvironmentDistribution): cells of the mean
// scaled radiance times sin(theta) at the row's centre; their densities
// over the mean, the rows' cumulative distribution and each row's. Returns
// the mean cell value (0 for a black image).
double environment_distribution(const lotus_hip_environment_map& e, uint32_t& columns, uint32_t& rows,
                                std::vector<float>& dist) {
  constexpr double kPiD = 3.14159265358979323846;
  columns = std::min(e.width, kMaxEnvironmentColumns);
  rows = std::min(e.height, kMaxEnvironmentRows);
  const size_t cells = size_t(columns) * rows;
  std::vector<double> sums(cells, 0.0);
  std::vector<uint32_t> counts(cells, 0);
  for (uint32_t y = 0; y < e.height; ++y) {
    const size_t row = uint64_t(y) * rows / e.height;
    for (uint32_t x = 0; x < e.width; ++x) {
      const size_t column = uint64_t(x) * columns / e.width;
      const float* t = static_cast<const float*>(e.texels) + 4 * (size_t(y) * e.width + x);
      double radiance = 0.0;
      for (int c = 0; c < 3; ++c) radiance += std::max(double(t[c]), 0.0) * e.scale[c];
      sums[row * columns + column] += radiance / 3.0;
      ++counts[row * columns + column];
    }
  }
  std::vector<double> values(cells), row_sums(rows, 0.0);
  double total = 0.0;
  for (uint32_t row = 0; row < rows; ++row) {
    const double sin_theta = std::sin(kPiD * (row + 0.5) / rows);
    for (uint32_t column = 0; column < columns; ++column) {
      const size_t cell = size_t(row) * columns + column;
      values[cell] = sums[cell] / counts[cell] * sin_theta;
      row_sums[row] += values[cell];
    }
    total += row_sums[row];
  }
  const double mean = total / static_cast<double>(cells);
  dist.assign(cells + (rows + 1) + size_t(rows) * (columns + 1), 0.0f);
  if (!(mean > 0.0) || !std::isfinite(mean)) return 0.0;
  float* density = dist.data();
  float* marginal = density + cells;
  float* conditional = marginal + rows + 1;
  for (size_t cell = 0; cell < cells; ++cell) density[cell] = static_cast<float>(values[cell] / mean);
  double cumulative = 0.0;
  for (uint32_t row = 0; row < rows; ++row) {
    cumulative += row_sums[row];
    marginal[row + 1] = static_cast<float>(cumulative / total);
    float* cdf = conditional + size_t(row) * (columns + 1);
    double along = 0.0;
    for (uint32_t column = 0; column < columns; ++column) {
      along += values[size_t(row) * columns + column];
      cdf[column + 1] = row_sums[row] > 0.0 ? static_cast<float>(along / row_sums[row])
                                            : static_cast<float>(column + 1) / columns;
    }
    cdf[columns] = 1.0f;
  }
  marginal[rows] = 1.0f;
  return mean;
}

// gpu_scene.cpp:144-203, 512-609 (BuildLightTable): one entry per triangle
// of every instance whose material emits (mean emission x world area, in
// double), then one per analytic light of positive weight, chosen by Vose's
// alias method; the crossing hierarchy over the bounded lights and the
// distant lights' indices (gpu_scene.cpp:448-471).
void build_light_table(const lotus_hip_scene_plan* p, LightTableHost& t) {
  t = LightTableHost{};
  t.offsets.assign(p->instance_count, kNoLights);
  std::vector<double> weights;
  for (uint32_t i = 0; i < p->instance_count; ++i) {
    const uint32_t* m = p->materials + 36ull * p->instance_material[i];
    const double mean = emission_estimate(m, p->bindings);
    const uint32_t geo = p->instance_geometry[i];
    const uint32_t count = p->geometry_triangle_count[geo];
    if (mean <= 0.0 || count == 0) continue;
    t.offsets[i] = static_cast<uint32_t>(t.lights.size());
    const uint32_t* tri = p->triangles + 3ull * p->geometry_triangle_offset[geo];
    const float* pos = p->positions + 3ull * p->geometry_vertex_offset[geo];
    const float* tm = p->instance_transform + 16ull * i;
    for (uint32_t k = 0; k < count; ++k) {
      double c[3][3];
      for (int v = 0; v < 3; ++v) {
        const float* q = pos + 3ull * tri[3 * k + v];
        for (int row = 0; row < 3; ++row)
          c[v][row] = double(tm[row]) * q[0] + double(tm[4 + row]) * q[1] + double(tm[8 + row]) * q[2] + tm[12 + row];
      }
      double a[3], b[3];
      for (int row = 0; row < 3; ++row) {
        a[row] = c[1][row] - c[0][row];
        b[row] = c[2][row] - c[0][row];
      }
      const double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
      double area = 0.5 * std::sqrt(x * x + y * y + z * z);
      if (!std::isfinite(area)) area = 0.0;
      weights.push_back(mean * area);
      t.lights.push_back(LightRecord{0.0f, 1.0f, 0u, i});
    }
  }
  double env_mean = 0.0;
  if (p->environment_map) {
    const lotus_hip_environment_map& e = *p->environment_map;
    double axes[3][3];
    environment_frame(e.world_from_environment, axes);  // validated
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) t.env_frame[3 * a + c] = static_cast<float>(axes[a][c]);
    for (int c = 0; c < 3; ++c) t.env_scale[c] = e.scale[c];
    t.env_width = e.width;
    t.env_height = e.height;
    t.env_texels.resize(size_t(e.width) * e.height);
    std::memcpy(t.env_texels.data(), e.texels, t.env_texels.size() * sizeof(float4));
    env_mean = environment_distribution(e, t.env_columns, t.env_rows, t.env_dist);
  }
  bool any_distant = env_mean > 0.0;
  for (uint32_t l = 0; l < p->light_count; ++l) any_distant |= p->lights[l].shape == LOTUS_HIP_LIGHT_DISTANT;
  const double radius = any_distant ? scene_radius(p) : 0.0;
  std::vector<uint32_t> analytic_entries;
  for (uint32_t l = 0; l < p->light_count; ++l) {
    AnalyticRecord rec;
    double weight = 0.0;
    if (!make_analytic_light(p->lights[l], radius, rec, weight)) continue;
    if (weight > 0.0) {
      analytic_entries.push_back(static_cast<uint32_t>(t.lights.size()));
      weights.push_back(weight);
      t.lights.push_back(LightRecord{0.0f, 1.0f, 0u, kAnalyticLight | static_cast<uint32_t>(t.analytic.size())});
    }
    t.analytic.push_back(rec);
  }
  std::vector<LightBox> bounded;
  for (uint32_t a = 0; a < t.analytic.size(); ++a) {
    const uint32_t shape = t.analytic[a].shape;
    if (shape == kDistantLight) {
      t.distant.push_back(a);
    } else if (shape == kParallelLight) {
      if (t.parallel_count++ == 0) t.parallel = a + 1;
    } else {
      bounded.push_back(bound_light(t.analytic[a], a));
    }
  }
  if (bounded.size() == 1) {
    LightNode& root = t.nodes.emplace_back();
    store_bounds(root.low0, root.high0, bounded.data(), bounded.data() + 1, kLightLeaf | bounded[0].index);
    store_bounds(root.low1, root.high1, nullptr, nullptr, kNoLightChild);
  } else if (bounded.size() > 1) {
    build_light_node(bounded.data(), bounded.data() + bounded.size(), t.nodes);
  }
  // gpu_scene.cpp:502-510 (EnvironmentWeight): pi R^2 times the image's
  // integral over the sphere, 2 pi^2 times the mean cell value, over 2 pi
  double env_weight = 0.0;
  if (env_mean > 0.0) {
    constexpr double kPiD = 3.14159265358979323846;
    env_weight = kPiD * kPiD * radius * radius * env_mean;
    if (!std::isfinite(env_weight)) env_weight = std::numeric_limits<double>::max();
  }
  if (env_weight > 0.0) {
    weights.push_back(env_weight);
    t.lights.push_back(LightRecord{0.0f, 1.0f, 0u, kEnvironmentLight});
  }
  double total = 0.0;
  for (double w : weights) total += w;
  if (!(total > 0.0) || !std::isfinite(total)) {
    t.offsets.assign(p->instance_count, kNoLights);
    t.lights.clear();
    return;
  }
  const size_t count = t.lights.size();
  std::vector<double> scaled(count);
  std::vector<uint32_t> small, large;
  for (size_t i = 0; i < count; ++i) {
    t.lights[i].pdf = static_cast<float>(weights[i] / total);
    scaled[i] = weights[i] / total * static_cast<double>(count);
    (scaled[i] < 1.0 ? small : large).push_back(static_cast<uint32_t>(i));
  }
  while (!small.empty() && !large.empty()) {
    const uint32_t less = small.back();
    small.pop_back();
    const uint32_t more = large.back();
    t.lights[less].threshold = static_cast<float>(scaled[less]);
    t.lights[less].alias = more;
    scaled[more] -= 1.0 - scaled[less];
    if (scaled[more] < 1.0) {
      large.pop_back();
      small.push_back(more);
    }
  }
  for (uint32_t i : small) {
    t.lights[i].threshold = 1.0f;
    t.lights[i].alias = i;
  }
  for (uint32_t i : large) {
    t.lights[i].threshold = 1.0f;
    t.lights[i].alias = i;
  }
  // each analytic record keeps its entry's probability (gpu_scene.cpp:600-604)
  for (uint32_t e : analytic_entries) t.analytic[t.lights[e].instance & ~kAnalyticLight].radiance[3] = t.lights[e].pdf;
  if (env_weight > 0.0) t.env_selection = t.lights.back().pdf;
}

// The plan's per-corner data in one comparable form: texture shapes, texel
// bytes and bindings; per instance the first corner of the texture
// coordinates its material reads, of its geometry's shading normals and of
// its colour set (bit 31: one colour for the whole mesh), 0xFFFFFFFF for
// none; and the coordinates, normals and colours those name.
constexpr uint32_t kNoCorner = 0xFFFFFFFFu;
constexpr uint32_t kConstantColor = 0x80000000u;
struct TextureData {
  std::vector<uint32_t> shape, bindings, instance_texcoords, instance_normals, instance_colors;
  std::vector<uint8_t> texel_bytes;
  std::vector<float> texcoords, normals, colors;
};

// Per instance its first texture-coordinate, normal and colour corner.
void instance_corners(const lotus_hip_scene_plan* p, std::vector<uint32_t>& tc, std::vector<uint32_t>& nm,
                      std::vector<uint32_t>& col) {
  const uint32_t I = p->instance_count;
  tc.assign(I, kNoCorner);
  nm.assign(I, kNoCorner);
  col.assign(I, kNoCorner);
  for (uint32_t i = 0; i < I; ++i) {
    const uint32_t g = p->instance_geometry[i];
    if (p->instance_texcoord_offset) tc[i] = p->instance_texcoord_offset[i];
    else if (p->geometry_texcoord_offset) tc[i] = p->geometry_texcoord_offset[g];
    if (p->geometry_normal_offset) nm[i] = p->geometry_normal_offset[g];
    if (p->instance_color_offset && p->instance_color_offset[i] != kNoCorner)
      col[i] = p->instance_color_offset[i] |
               (p->instance_color_constant && p->instance_color_constant[i] ? kConstantColor : 0u);
  }
}

// The corners the instances name: one past the last texture coordinate,
// normal and colour entry read.
void corner_extents(const lotus_hip_scene_plan* p, const std::vector<uint32_t>& tc, const std::vector<uint32_t>& nm,
                    const std::vector<uint32_t>& col, uint64_t& tcs, uint64_t& nms, uint64_t& cols) {
  tcs = nms = cols = 0;
  for (uint32_t i = 0; i < p->instance_count; ++i) {
    const uint64_t corners = 3ull * p->geometry_triangle_count[p->instance_geometry[i]];
    if (tc[i] != kNoCorner) tcs = std::max<uint64_t>(tcs, uint64_t(tc[i]) + corners);
    if (nm[i] != kNoCorner) nms = std::max<uint64_t>(nms, uint64_t(nm[i]) + corners);
    if (col[i] != kNoCorner)
      cols = std::max<uint64_t>(cols, uint64_t(col[i] & ~kConstantColor) + ((col[i] & kConstantColor) ? 1 : corners));
  }
}

bool material_emits(const uint32_t* m, const uint32_t* bindings);

// `trusted`: the LOTUS_HIP_UNCHANGED_* bits honoured (the plan's, when a
// scene is resident); vouched parts are not validated again.
std::string validate_plan(const lotus_hip_scene_plan* p, uint32_t trusted = 0) {
  if (!p) return "no scene plan";
  if (p->material_count == 0 || !p->materials) return "the material table is empty";
  if (p->bvh_leaf_size > 8) return "BVH leaves hold at most 8 triangles";
  if (p->bvh_builder > LOTUS_HIP_BVH_SAH) return "unknown BVH builder";
  if ((trusted & LOTUS_HIP_UNCHANGED_POSITIONS) && p->moved_geometry_count != 0 && !p->moved_geometries)
    return "the plan names moved geometries but holds none";
  for (uint32_t k = 0; (trusted & LOTUS_HIP_UNCHANGED_POSITIONS) && k < p->moved_geometry_count; ++k)
    if (p->moved_geometries[k] >= p->geometry_count) return "a moved geometry is not in the plan";
  const auto check_positions = [&](uint32_t g) -> std::string {
    const float* pos = p->positions + 3ull * p->geometry_vertex_offset[g];
    for (uint64_t k = 0; k < 3ull * p->geometry_vertex_count[g]; ++k)
      if (!std::isfinite(pos[k])) return "geometry " + std::to_string(g) + " has a non-finite position";
    return "";
  };
  uint64_t tris = 0;
  for (uint32_t g = 0; g < p->geometry_count; ++g) {
    const uint32_t vc = p->geometry_vertex_count[g];
    const uint32_t* tri = p->triangles + 3ull * p->geometry_triangle_offset[g];
    for (uint64_t k = 0; !(trusted & LOTUS_HIP_UNCHANGED_TOPOLOGY) && k < 3ull * p->geometry_triangle_count[g]; ++k)
      if (tri[k] >= vc) return "geometry " + std::to_string(g) + " indexes a missing vertex";
    if (!(trusted & LOTUS_HIP_UNCHANGED_POSITIONS)) {
      const std::string bad = check_positions(g);
      if (!bad.empty()) return bad;
    }
  }
  for (uint32_t k = 0; (trusted & LOTUS_HIP_UNCHANGED_POSITIONS) && k < p->moved_geometry_count; ++k) {
    const std::string bad = check_positions(p->moved_geometries[k]);
    if (!bad.empty()) return bad;
  }

  for (uint32_t i = 0; i < p->instance_count; ++i) {
    if (p->instance_geometry[i] >= p->geometry_count) return "an instance names a missing geometry";
    if (p->instance_material[i] >= p->material_count) return "an instance names a missing material slot";
    for (int k = 0; k < 16; ++k)
      if (!std::isfinite(p->instance_transform[16ull * i + k])) return "an instance transform is not finite";
    tris += p->geometry_triangle_count[p->instance_geometry[i]];
  }
  if (tris >= (1ull << 27)) return "the scene holds more than 2^27 triangles";
  if (p->light_count != 0 && !p->lights) return "the plan names lights but holds none";
  if ((p->texture_count != 0 && !p->textures) || (p->binding_count != 0 && !p->bindings))
    return "the plan names textures or bindings but holds none";
  uint64_t texel_total = 0;
  for (uint32_t k = 0; k < p->texture_count; ++k) {
    const lotus_hip_texture& x = p->textures[k];
    if (x.width == 0 || x.height == 0 || x.width > 16384 || x.height > 16384 || !x.texels ||
        x.format > LOTUS_HIP_TEXTURE_RGBA32F)
      return "texture " + std::to_string(k) + " must be 1 to 16384 texels on a side in a known format";
    texel_total += uint64_t(x.width) * x.height;
  }
  // each texture's first texel is a 32-bit offset into the shared texel array
  if (texel_total > 0xFFFFFFFFull) return "the plan's textures hold more than 2^32 - 1 texels";
  for (uint32_t k = 0; k < p->binding_count; ++k) {
    const uint32_t* b = p->bindings + 24ull * k;
    if (b[3] > 2u || b[2] > 3u || b[1] > 15u || (b[3] == 1u && b[0] >= p->texture_count))
      return "binding " + std::to_string(k) + " names a missing texture or an unknown mode";
    for (int w = 4; w < 24; ++w) {
      float f;
      std::memcpy(&f, b + w, 4);
      if (!std::isfinite(f)) return "binding " + std::to_string(k) + " holds a non-finite value";
    }
  }
  uint64_t texels = 0;
  for (uint32_t k = 0; k < p->texture_count; ++k) texels += uint64_t(p->textures[k].width) * p->textures[k].height;
  if (texels >= (1ull << 32)) return "the textures hold 2^32 texels or more";
  {
    std::vector<uint32_t> tc, nm, col;
    instance_corners(p, tc, nm, col);
    uint64_t tcs, nms, cols;
    corner_extents(p, tc, nm, col, tcs, nms, cols);
    if (tcs && !p->texcoords) return "the plan names texture coordinates but holds none";
    if (nms && !p->normals) return "the plan names shading normals but holds none";
    if (cols && !p->colors) return "the plan names colours but holds none";
    if (tcs >= (1ull << 32) || nms >= (1ull << 32) || cols >= (1ull << 31))
      return "the plan's per-corner data holds too many corners";
    for (uint64_t k = 0; k < 2 * tcs; ++k)
      if (!std::isfinite(p->texcoords[k])) return "the plan has a non-finite texture coordinate";
    if (p->instance_color_offset)
      for (uint32_t i = 0; i < p->instance_count; ++i)
        if (p->instance_color_offset[i] != kNoCorner && p->instance_color_offset[i] >= kConstantColor)
          return "a colour set's first entry lies beyond 2^31";
  }
  if (p->environment_map) {
    const lotus_hip_environment_map& e = *p->environment_map;
    if (e.width == 0 || e.height == 0 || e.width > 16384 || e.height > 16384 || !e.texels)
      return "the environment image must be 1 to 16384 texels on a side";
    if (e.format != LOTUS_HIP_TEXTURE_RGBA32F) return "unknown environment texel format";
    const float* texels = static_cast<const float*>(e.texels);
    for (size_t k = 0; k < 4ull * e.width * e.height; ++k)
      if (!std::isfinite(texels[k])) return "the environment image has a non-finite texel";
    for (int c = 0; c < 3; ++c)
      if (!(e.scale[c] >= 0.0f) || !std::isfinite(e.scale[c])) return "the environment scale must be finite and non-negative";
    double axes[3][3];
    if (!environment_frame(e.world_from_environment, axes)) return "the environment frame's axes are not independent";
  }
  if (p->light_count > (1u << 24)) return "the scene holds more than 2^24 analytic lights";
  for (uint32_t l = 0; l < p->light_count; ++l) {
    const lotus_hip_light& L = p->lights[l];
    const std::string which = "light " + std::to_string(l);
    if (L.shape > LOTUS_HIP_LIGHT_DISTANT) return which + " has an unknown shape";
    for (int k = 0; k < 16; ++k)
      if (!std::isfinite(L.world_from_light[k])) return which + ": its transform is not finite";
    for (int c = 0; c < 3; ++c)
      if (!(L.radiance[c] >= 0.0f) || !std::isfinite(L.radiance[c])) return which + ": radiance must be finite and non-negative";
    if (!(L.radius >= 0.0f) || !(L.width >= 0.0f) || !(L.height >= 0.0f) || !(L.angle >= 0.0f) ||
        !std::isfinite(L.radius) || !std::isfinite(L.width) || !std::isfinite(L.height) || !std::isfinite(L.angle))
      return which + ": sizes must be finite and non-negative";
  }
  for (uint32_t m = 0; m < p->material_count; ++m) {
    const uint32_t* w = p->materials + 36ull * m;
    const uint32_t model = w[0] & 0xFFu;
    if (model > 1u) return "material " + std::to_string(m) + ": only UsdPreviewSurface and glTF PBR are supported";
    // UsdPreviewSurface reads slots 0-6, glTF PBR slots 0-13 (scene.slang:59-79)
    if (w[3] != 0 || (w[2] & ~(model == 1u ? 0x3FFFu : 0x7Fu)) != 0)
      return "material " + std::to_string(m) + ": a lookup in a slot its model does not read";
    if (model == 1u) {
      // ValidGltfPbr (gpu_scene.cpp:853-881)
      float f[36];
      std::memcpy(f, w, sizeof f);
      auto unit = [](float x) { return x >= 0.0f && x <= 1.0f; };
      auto radiance = [](float x) { return std::isfinite(x) && x >= 0.0f; };
      const bool ok = unit(f[12]) && unit(f[13]) && unit(f[14]) && unit(f[15]) && unit(f[19]) && radiance(f[16]) &&
                      radiance(f[17]) && radiance(f[18]) && radiance(f[20]) && unit(f[21]) && unit(f[22]) &&
                      unit(f[23]) && radiance(f[24]) && radiance(f[25]) && radiance(f[26]) && unit(f[27]) &&
                      unit(f[28]) && unit(f[29]) && unit(f[30]) && unit(f[31]) && unit(f[32]) &&
                      std::isfinite(f[33]) && f[33] >= 1.0f;
      if (!ok) return "material " + std::to_string(m) + ": a glTF PBR parameter lies outside its range";
    }
    if (w[2] != 0 && (!p->bindings || uint64_t(w[1]) + __builtin_popcount(w[2]) > p->binding_count))
      return "material " + std::to_string(m) + ": its lookups lie outside the binding table";
    float opacity, threshold;
    std::memcpy(&opacity, w + 6, 4);
    std::memcpy(&threshold, w + 7, 4);
    // ValidGeometry (gpu_scene.cpp:821-831)
    if (!(opacity >= 0.0f && opacity <= 1.0f) || !(threshold >= 0.0f && threshold <= 1.0f))
      return "material " + std::to_string(m) + ": opacity and its threshold must lie in [0, 1]";
  }
  // with the topology vouched for, the light table still reads the emitting
  // instances' triangles from the plan: their indices are checked
  if (trusted & LOTUS_HIP_UNCHANGED_TOPOLOGY) {
    for (uint32_t i = 0; i < p->instance_count; ++i) {
      const uint32_t g = p->instance_geometry[i];
      if (!material_emits(p->materials + 36ull * p->instance_material[i], p->bindings)) continue;
      const uint32_t vc = p->geometry_vertex_count[g];
      const uint32_t* tri = p->triangles + 3ull * p->geometry_triangle_offset[g];
      for (uint64_t k = 0; k < 3ull * p->geometry_triangle_count[g]; ++k)
        if (tri[k] >= vc) return "geometry " + std::to_string(g) + " indexes a missing vertex";
    }
  }
  return {};
}

// The plan's materials of the glTF PBR model.
uint32_t count_gltf_materials(const lotus_hip_scene_plan* p) {
  uint32_t n = 0;
  for (uint32_t m = 0; m < p->material_count; ++m) n += (p->materials[36ull * m] & 0xFFu) == 1u ? 1u : 0u;
  return n;
}

// Per instance 1 when a ray tests its material's coverage (scene.slang:
// 393-406): an opacity lookup, or a constant coverage below one (opacity
// below one without a threshold, or below a positive threshold).
std::vector<uint32_t> covered_flags(const lotus_hip_scene_plan* p) {
  std::vector<uint32_t> covered(p->instance_count, 0u);
  for (uint32_t i = 0; i < p->instance_count; ++i) {
    const uint32_t* w = p->materials + 36ull * p->instance_material[i];
    float opacity, threshold;
    std::memcpy(&opacity, w + 6, 4);
    std::memcpy(&threshold, w + 7, 4);
    const float constant = threshold > 0.0f ? (opacity >= threshold ? 1.0f : 0.0f) : opacity;
    covered[i] = ((w[2] & 1u) != 0 || !(constant >= 1.0f)) ? 1u : 0u;
  }
  return covered;
}

// The extents of a plan's arrays.
struct PlanShape {
  uint32_t geometries = 0, instances = 0, materials = 0, tri_count = 0, leaf_size = 2, builder = 0;
  uint64_t vertex_words = 0, index_words = 0;
  std::vector<uint32_t> base;  // first flattened triangle per instance, instances + 1
};

PlanShape plan_shape(const lotus_hip_scene_plan* p) {
  PlanShape sh;
  sh.geometries = p->geometry_count;
  sh.instances = p->instance_count;
  sh.materials = p->material_count;
  sh.leaf_size = p->bvh_leaf_size ? p->bvh_leaf_size : 2u;
  sh.builder = p->bvh_builder;
  sh.base.assign(sh.instances + 1, 0);
  for (uint32_t i = 0; i < sh.instances; ++i)
    sh.base[i + 1] = sh.base[i] + p->geometry_triangle_count[p->instance_geometry[i]];
  sh.tri_count = sh.base[sh.instances];
  for (uint32_t g = 0; g < sh.geometries; ++g) {
    sh.vertex_words =
        std::max<uint64_t>(sh.vertex_words, 3ull * (uint64_t(p->geometry_vertex_offset[g]) + p->geometry_vertex_count[g]));
    sh.index_words = std::max<uint64_t>(
        sh.index_words, 3ull * (uint64_t(p->geometry_triangle_offset[g]) + p->geometry_triangle_count[g]));
  }
  return sh;
}

template <typename T>
bool same_words(const std::vector<T>& held, const T* src, uint64_t n) {
  return held.size() == n && (n == 0 || std::memcmp(held.data(), src, n * sizeof(T)) == 0);
}

// An instance of this material enters the light table (gpu_scene.cpp:144-203).
bool material_emits(const uint32_t* m, const uint32_t* bindings) { return emission_estimate(m, bindings) > 0.0; }

TextureData texture_data(const lotus_hip_scene_plan* p) {
  TextureData d;
  for (uint32_t k = 0; k < p->texture_count; ++k) {
    const lotus_hip_texture& x = p->textures[k];
    d.shape.insert(d.shape.end(), {x.width, x.height, x.format});
    const size_t bytes = size_t(x.width) * x.height * (x.format == LOTUS_HIP_TEXTURE_RGBA32F ? 16 : 4);
    const uint8_t* t = static_cast<const uint8_t*>(x.texels);
    d.texel_bytes.insert(d.texel_bytes.end(), t, t + bytes);
  }
  if (p->binding_count) d.bindings.assign(p->bindings, p->bindings + 24ull * p->binding_count);
  instance_corners(p, d.instance_texcoords, d.instance_normals, d.instance_colors);
  uint64_t tcs, nms, cols;
  corner_extents(p, d.instance_texcoords, d.instance_normals, d.instance_colors, tcs, nms, cols);
  if (tcs) d.texcoords.assign(p->texcoords, p->texcoords + 2 * tcs);
  if (nms) d.normals.assign(p->normals, p->normals + 3 * nms);
  if (cols) d.colors.assign(p->colors, p->colors + 3 * cols);
  return d;
}

// What a plan changes against the applied one (GpuScene::Apply's retain and
// TLAS-current tests, gpu_scene.cpp:1638-1725).
struct Changes {
  uint32_t update = LOTUS_HIP_SCENE_BUILD;
  std::vector<uint32_t> moved;  // geometries whose positions differ
  bool transforms = false, materials = false, instance_materials = false, environment = false, lights = false;
  bool refit = false;
  bool temporal_compatible = false;
};

Changes compare_plan(const lotus_hip_renderer* r, const lotus_hip_scene_plan* p, const PlanShape& sh) {
  Changes c;
  const auto& old = r->plan;
  const uint32_t u = r->has_scene ? p->unchanged : 0u;  // vouched-for parts
  // A retained geometry keeps its vertex count and exactly its triangle
  // indices (gpu_scene.cpp:1651-1656); every instance keeps its geometry.
  const bool shapes = r->has_scene && old.leaf_size == sh.leaf_size && old.builder == sh.builder &&
                      old.gvo.size() == sh.geometries && old.igeo.size() == sh.instances &&
                      old.positions.size() == sh.vertex_words && old.triangles.size() == sh.index_words;
  const bool topology =
      shapes && ((u & LOTUS_HIP_UNCHANGED_TOPOLOGY) ||
                 (same_words(old.gvo, p->geometry_vertex_offset, sh.geometries) &&
                  same_words(old.gvc, p->geometry_vertex_count, sh.geometries) &&
                  same_words(old.gto, p->geometry_triangle_offset, sh.geometries) &&
                  same_words(old.gtc, p->geometry_triangle_count, sh.geometries) &&
                  same_words(old.igeo, p->instance_geometry, sh.instances) &&
                  same_words(old.triangles, p->triangles, sh.index_words)));
  if (!topology) return c;
  // texture lookups are uploaded with a new scene (a change of them builds)
  if (!(u & LOTUS_HIP_UNCHANGED_TOPOLOGY)) {
    const TextureData tex = texture_data(p);
    if (tex.shape != old.texture_shape || tex.bindings != old.bindings ||
        tex.instance_texcoords != old.instance_texcoords || tex.instance_normals != old.instance_normals ||
        tex.instance_colors != old.instance_colors || tex.texel_bytes != old.texel_bytes ||
        tex.texcoords != old.texcoords || tex.normals != old.normals || tex.colors != old.colors)
      return c;
  }
  if (u & LOTUS_HIP_UNCHANGED_POSITIONS) {
    std::vector<char> seen(sh.geometries, 0);
    for (uint32_t k = 0; k < p->moved_geometry_count; ++k) {
      const uint32_t g = p->moved_geometries[k];
      if (seen[g]) continue;
      seen[g] = 1;
      // a moved geometry's positions are uploaded at its place: that place
      // must be the resident one, or the plan builds (validated whole)
      if (p->geometry_vertex_offset[g] != old.gvo[g] || p->geometry_vertex_count[g] != old.gvc[g]) {
        c.moved.clear();
        return c;
      }
      const uint64_t first = 3ull * p->geometry_vertex_offset[g], words = 3ull * p->geometry_vertex_count[g];
      if (words && std::memcmp(old.positions.data() + first, p->positions + first, words * 4) != 0) c.moved.push_back(g);
    }
    std::sort(c.moved.begin(), c.moved.end());
  } else if (!same_words(old.positions, p->positions, sh.vertex_words)) {
    for (uint32_t g = 0; g < sh.geometries; ++g) {
      const uint64_t first = 3ull * p->geometry_vertex_offset[g], words = 3ull * p->geometry_vertex_count[g];
      if (words && std::memcmp(old.positions.data() + first, p->positions + first, words * 4) != 0) c.moved.push_back(g);
    }
  }
  c.transforms = !(u & LOTUS_HIP_UNCHANGED_TRANSFORMS) &&
                 !same_words(old.transforms, p->instance_transform, 16ull * sh.instances);
  c.materials = !(u & LOTUS_HIP_UNCHANGED_MATERIALS) && !same_words(old.materials, p->materials, 36ull * sh.materials);
  c.instance_materials =
      !(u & LOTUS_HIP_UNCHANGED_MATERIALS) && !same_words(old.imat, p->instance_material, sh.instances);
  const bool lights_vouched = (u & LOTUS_HIP_UNCHANGED_LIGHTS) != 0;
  c.environment = !lights_vouched && std::memcmp(old.environment, p->environment, sizeof old.environment) != 0;
  c.lights = !lights_vouched && !same_words(old.lights, p->lights, p->light_count);
  const lotus_hip_environment_map* e = p->environment_map;
  const bool env_same =
      lights_vouched || old.has_env == (e != nullptr) &&
      (!e || (old.env_width == e->width && old.env_height == e->height &&
              std::memcmp(old.env_scale, e->scale, sizeof old.env_scale) == 0 &&
              std::memcmp(old.env_frame, e->world_from_environment, sizeof old.env_frame) == 0 &&
              same_words(old.env_texels, static_cast<const float*>(e->texels), 4ull * e->width * e->height)));
  c.lights = c.lights || !env_same;
  c.refit = !c.moved.empty() || c.transforms;
  if (!c.refit && !c.materials && !c.instance_materials && !c.environment && !c.lights) {
    c.update = LOTUS_HIP_SCENE_UNCHANGED;
    c.temporal_compatible = true;
    return c;
  }
  c.update = c.refit ? LOTUS_HIP_SCENE_REFIT : LOTUS_HIP_SCENE_TABLES;
  // TemporalSceneCompatible (vulkan_backend.cpp:262-289): the environment,
  // the material table and the instances' slots are kept, and an emitting
  // instance keeps its transform and its geometry's positions.
  c.temporal_compatible = !c.materials && !c.instance_materials && !c.environment && !c.lights;
  if (c.temporal_compatible && c.refit) {
    std::vector<char> moved(sh.geometries, 0);
    for (uint32_t g : c.moved) moved[g] = 1;
    for (uint32_t i = 0; i < sh.instances && c.temporal_compatible; ++i) {
      if (!material_emits(p->materials + 36ull * p->instance_material[i], p->bindings)) continue;
      if (moved[p->instance_geometry[i]] ||
          std::memcmp(old.transforms.data() + 16ull * i, p->instance_transform + 16ull * i, 64) != 0)
        c.temporal_compatible = false;
    }
  }
  return c;
}

void record_plan(lotus_hip_renderer* r, const lotus_hip_scene_plan* p, const PlanShape& sh, const Changes& c) {
  auto& d = r->plan;
  const bool all = c.update == LOTUS_HIP_SCENE_BUILD;
  if (all) {
    TextureData tex = texture_data(p);
    d.texture_shape = std::move(tex.shape);
    d.bindings = std::move(tex.bindings);
    d.instance_texcoords = std::move(tex.instance_texcoords);
    d.instance_normals = std::move(tex.instance_normals);
    d.instance_colors = std::move(tex.instance_colors);
    d.texel_bytes = std::move(tex.texel_bytes);
    d.texcoords = std::move(tex.texcoords);
    d.normals = std::move(tex.normals);
    d.colors = std::move(tex.colors);
    d.positions.assign(p->positions, p->positions + sh.vertex_words);
    d.triangles.assign(p->triangles, p->triangles + sh.index_words);
    d.gvo.assign(p->geometry_vertex_offset, p->geometry_vertex_offset + sh.geometries);
    d.gvc.assign(p->geometry_vertex_count, p->geometry_vertex_count + sh.geometries);
    d.gto.assign(p->geometry_triangle_offset, p->geometry_triangle_offset + sh.geometries);
    d.gtc.assign(p->geometry_triangle_count, p->geometry_triangle_count + sh.geometries);
    d.igeo.assign(p->instance_geometry, p->instance_geometry + sh.instances);
  } else {
    for (uint32_t g : c.moved) {
      const uint64_t first = 3ull * p->geometry_vertex_offset[g], words = 3ull * p->geometry_vertex_count[g];
      std::memcpy(d.positions.data() + first, p->positions + first, words * 4);
    }
  }
  if (all || c.transforms) d.transforms.assign(p->instance_transform, p->instance_transform + 16ull * sh.instances);
  if (all || c.materials) d.materials.assign(p->materials, p->materials + 36ull * sh.materials);
  if (all || c.lights) {
    d.lights.assign(p->lights, p->lights + p->light_count);
    const lotus_hip_environment_map* e = p->environment_map;
    d.has_env = e != nullptr;
    d.env_width = e ? e->width : 0;
    d.env_height = e ? e->height : 0;
    if (e) {
      const float* texels = static_cast<const float*>(e->texels);
      d.env_texels.assign(texels, texels + 4ull * e->width * e->height);
      std::memcpy(d.env_scale, e->scale, sizeof d.env_scale);
      std::memcpy(d.env_frame, e->world_from_environment, sizeof d.env_frame);
    } else {
      d.env_texels.clear();
    }
  }
  if (all || c.instance_materials) d.imat.assign(p->instance_material, p->instance_material + sh.instances);
  std::memcpy(d.environment, p->environment, sizeof d.environment);
  d.leaf_size = sh.leaf_size;
  d.builder = sh.builder;
}

void set_view(lotus_hip_renderer* r, const PlanShape& sh, const float* environment) {
  const SceneBuffers& s = r->scene;
  SceneView& v = r->view;
  v.nodes = s.nodes;
  v.tris = s.tris;
  v.tri_of_global = s.tri_of_global;
  v.instance_base = s.instance_base;
  v.instance_material = s.instance_material;
  v.instance_transform = s.instance_transform;
  v.materials = s.materials;
  v.light_offsets = s.light_offsets;
  v.lights = s.lights;
  v.light_count = s.light_count;
  v.analytic = s.analytic;
  v.light_nodes = s.light_nodes;
  v.distant = s.distant;
  v.analytic_count = s.analytic_count;
  v.light_node_count = s.light_node_count;
  v.distant_count = s.distant_count;
  v.parallel = s.parallel;
  v.parallel_count = s.parallel_count;
  v.gltf_materials = s.gltf_materials;
  v.albedo_tables = r->albedo_tables;
  v.env_texels = s.env_texels;
  v.env_dist = s.env_dist;
  v.env_width = s.env_width;
  v.env_height = s.env_height;
  v.env_columns = s.env_columns;
  v.env_rows = s.env_rows;
  std::memcpy(v.env_frame, s.env_frame, sizeof v.env_frame);
  std::memcpy(v.env_scale, s.env_scale, sizeof v.env_scale);
  v.env_selection = s.env_selection;
  v.textures = s.textures;
  v.texels = s.texels;
  v.bindings = s.binding_count ? s.bindings : nullptr;
  v.texcoords = s.texcoord_count ? s.texcoords : nullptr;
  v.instance_texcoords = s.instance_texcoords;
  v.normals = s.normals;
  v.instance_normals = s.normal_count ? s.instance_normals : nullptr;
  v.colors = s.colors;
  v.instance_colors = s.color_count ? s.instance_colors : nullptr;
  v.covered_instances = s.covered_instances;
  v.tri_count = sh.tri_count;
  v.instance_count = sh.instances;
  for (int c = 0; c < 3; ++c) v.environment[c] = environment[c];
}

uint64_t scene_device_bytes(const lotus_hip_renderer* r) {
  const SceneBuffers& s = r->scene;
  const uint64_t I = r->view.instance_count, G = r->plan.gvo.size();
  return s.bytes + 4 * (s.vertex_words + s.index_words) + 4 * (2 * G + I) + 4 * (I + 1) + 68 * I +
         144ull * s.material_count + 4 * I + uint64_t(s.light_count) * sizeof(LightRecord) +
         uint64_t(s.analytic_count) * sizeof(AnalyticRecord) + uint64_t(s.light_node_count) * sizeof(LightNode) +
         4ull * s.distant_count + 16ull * s.env_width * s.env_height +
         4ull * (s.env_columns == 0 ? 0 : uint64_t(s.env_columns) * s.env_rows + s.env_rows + 1 +
                                              uint64_t(s.env_rows) * (s.env_columns + 1)) +
         16ull * (s.texture_count + s.texel_count) + 96ull * s.binding_count + 8ull * s.texcoord_count + 4ull * I +
         12ull * (s.normal_count + s.color_count) + 8ull * I;
}

void fill_evidence(const lotus_hip_renderer* r, lotus_hip_scene_evidence* ev, uint32_t update, uint32_t moved,
                   double upload_ms, double build_ms) {
  if (!ev) return;
  const SceneBuffers& s = r->scene;
  ev->triangle_count = r->view.tri_count;
  ev->instance_count = r->view.instance_count;
  ev->material_count = s.material_count;
  ev->light_count = s.light_count;
  ev->node_count = s.node_count;
  ev->bvh_depth = s.depth;
  ev->leaf_count = s.leaf_count;
  ev->bvh_width = kWidth;
  ev->upload_ms = upload_ms;
  ev->build_gpu_ms = build_ms;
  ev->device_bytes = scene_device_bytes(r);
  ev->update = update;
  ev->moved_geometries = moved;
  ev->bvh_builds = r->bvh_builds;
  ev->bvh_refits = r->bvh_refits;
  ev->material_writes = r->material_writes;
  ev->light_writes = r->light_writes;
  ev->uploaded_bytes = r->uploaded_bytes;
}

// The light table's device buffers: sized to `t` (reallocated when a size
// changes) and filled; `up` enqueues a copy.
template <typename Up>
hipError_t upload_light_table(SceneBuffers& s, const LightTableHost& t, Up&& up) {
  hipError_t e = hipSuccess;
  auto fit = [&](auto*& ptr, size_t have, size_t need) {
    if (e == hipSuccess && (ptr == nullptr || have != need)) {
      dev_free(ptr);
      e = dev_alloc(ptr, need);
    }
  };
  fit(s.lights, s.light_count, t.lights.size());
  fit(s.analytic, s.analytic_count, t.analytic.size());
  fit(s.light_nodes, s.light_node_count, t.nodes.size());
  fit(s.distant, s.distant_count, t.distant.size());
  fit(s.env_texels, size_t(s.env_width) * s.env_height, t.env_texels.size());
  fit(s.env_dist, s.env_columns == 0 ? 0 : size_t(s.env_columns) * s.env_rows + s.env_rows + 1 +
                                               size_t(s.env_rows) * (s.env_columns + 1),
      t.env_dist.size());
  if (e != hipSuccess) return e;
  s.env_width = t.env_width;
  s.env_height = t.env_height;
  s.env_columns = t.env_columns;
  s.env_rows = t.env_rows;
  std::memcpy(s.env_frame, t.env_frame, sizeof s.env_frame);
  std::memcpy(s.env_scale, t.env_scale, sizeof s.env_scale);
  s.env_selection = t.env_selection;
  up(s.env_texels, t.env_texels.data(), t.env_texels.size() * sizeof(float4));
  up(s.env_dist, t.env_dist.data(), t.env_dist.size() * 4);
  s.light_count = static_cast<uint32_t>(t.lights.size());
  s.analytic_count = static_cast<uint32_t>(t.analytic.size());
  s.light_node_count = static_cast<uint32_t>(t.nodes.size());
  s.distant_count = static_cast<uint32_t>(t.distant.size());
  s.parallel_count = t.parallel_count;
  s.parallel = t.parallel;
  up(s.light_offsets, t.offsets.data(), t.offsets.size() * 4);
  up(s.lights, t.lights.data(), t.lights.size() * sizeof(LightRecord));
  up(s.analytic, t.analytic.data(), t.analytic.size() * sizeof(AnalyticRecord));
  up(s.light_nodes, t.nodes.data(), t.nodes.size() * sizeof(LightNode));
  up(s.distant, t.distant.data(), t.distant.size() * 4);
  return hipSuccess;
}

// A new scene: every buffer uploaded, a new BVH built.
bool apply_build(lotus_hip_renderer* r, const lotus_hip_scene_plan* p, const PlanShape& sh, double& upload_ms,
                 double& build_ms, std::string& error) {
  const auto t0 = std::chrono::steady_clock::now();
  LightTableHost table;
  build_light_table(p, table);
  free_scene(r);
  SceneBuffers& s = r->scene;
  const uint32_t G = sh.geometries, I = sh.instances, M = sh.materials;
  hipError_t e = dev_alloc(s.positions, sh.vertex_words);
  if (e == hipSuccess) e = dev_alloc(s.triangles, sh.index_words);
  if (e == hipSuccess) e = dev_alloc(s.geometry_vertex_offset, G);
  if (e == hipSuccess) e = dev_alloc(s.geometry_triangle_offset, G);
  if (e == hipSuccess) e = dev_alloc(s.instance_geometry, I);
  if (e == hipSuccess) e = dev_alloc(s.instance_base, I + 1);
  if (e == hipSuccess) e = dev_alloc(s.instance_material, I);
  if (e == hipSuccess) e = dev_alloc(s.instance_covered, I);
  if (e == hipSuccess) e = dev_alloc(s.instance_transform, 16ull * I);
  if (e == hipSuccess) e = dev_alloc(s.materials, 36ull * M);
  if (e == hipSuccess) e = dev_alloc(s.light_offsets, I);
  uint64_t bytes = 0;
  auto up = [&](void* dst, const void* src, size_t n) {
    if (e == hipSuccess && n) {
      e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, r->stream);
      bytes += n;
    }
  };
  up(s.positions, p->positions, sh.vertex_words * 4);
  up(s.triangles, p->triangles, sh.index_words * 4);
  up(s.geometry_vertex_offset, p->geometry_vertex_offset, G * 4ull);
  up(s.geometry_triangle_offset, p->geometry_triangle_offset, G * 4ull);
  up(s.instance_geometry, p->instance_geometry, I * 4ull);
  up(s.instance_base, sh.base.data(), sh.base.size() * 4);
  up(s.instance_material, p->instance_material, I * 4ull);
  s.covered_host = covered_flags(p);
  s.covered_instances = static_cast<uint32_t>(std::count(s.covered_host.begin(), s.covered_host.end(), 1u));
  up(s.instance_covered, s.covered_host.data(), I * 4ull);
  up(s.instance_transform, p->instance_transform, 64ull * I);
  up(s.materials, p->materials, 144ull * M);
  s.gltf_materials = count_gltf_materials(p);
  if (e == hipSuccess) e = upload_light_table(s, table, up);
  // texture lookups: texels decoded to linear RGBA, {first texel, width,
  // height} per texture, the binding table, texture coordinates
  std::vector<float4> texels;
  std::vector<uint4> textures;
  for (uint32_t k = 0; k < p->texture_count; ++k) {
    const lotus_hip_texture& x = p->textures[k];
    textures.push_back(make_uint4(static_cast<uint32_t>(texels.size()), x.width, x.height, 0u));
    const size_t n = size_t(x.width) * x.height;
    if (x.format == LOTUS_HIP_TEXTURE_RGBA32F) {
      const float4* f = static_cast<const float4*>(x.texels);
      texels.insert(texels.end(), f, f + n);
      continue;
    }
    const uint8_t* b = static_cast<const uint8_t*>(x.texels);
    for (size_t i = 0; i < n; ++i) {
      float c[4];
      for (int k2 = 0; k2 < 4; ++k2) {
        double v = b[4 * i + k2] / 255.0;
        // sRGB texels decode their colour, not alpha (TexelRgb, gpu_scene.cpp:612-629)
        if (x.format == LOTUS_HIP_TEXTURE_RGBA8_SRGB && k2 < 3)
          v = v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4);
        c[k2] = static_cast<float>(v);
      }
      texels.push_back(make_float4(c[0], c[1], c[2], c[3]));
    }
  }
  const TextureData tex = texture_data(p);
  s.texture_count = p->texture_count;
  s.binding_count = p->binding_count;
  s.texel_count = texels.size();
  s.texcoord_count = tex.texcoords.size() / 2;
  s.normal_count = tex.normals.size() / 3;
  s.color_count = tex.colors.size() / 3;
  if (e == hipSuccess) e = dev_alloc(s.textures, textures.size());
  if (e == hipSuccess) e = dev_alloc(s.texels, texels.size());
  if (e == hipSuccess) e = dev_alloc(s.bindings, tex.bindings.size());
  if (e == hipSuccess) e = dev_alloc(s.texcoords, tex.texcoords.size() / 2);
  if (e == hipSuccess) e = dev_alloc(s.instance_texcoords, I);
  if (e == hipSuccess) e = dev_alloc(s.normals, tex.normals.size());
  if (e == hipSuccess) e = dev_alloc(s.instance_normals, I);
  if (e == hipSuccess) e = dev_alloc(s.colors, tex.colors.size());
  if (e == hipSuccess) e = dev_alloc(s.instance_colors, I);
  up(s.textures, textures.data(), textures.size() * sizeof(uint4));
  up(s.texels, texels.data(), texels.size() * sizeof(float4));
  up(s.bindings, tex.bindings.data(), tex.bindings.size() * 4);
  up(s.texcoords, tex.texcoords.data(), tex.texcoords.size() * 4);
  up(s.instance_texcoords, tex.instance_texcoords.data(), tex.instance_texcoords.size() * 4);
  up(s.normals, tex.normals.data(), tex.normals.size() * 4);
  up(s.instance_normals, tex.instance_normals.data(), tex.instance_normals.size() * 4);
  up(s.colors, tex.colors.data(), tex.colors.size() * 4);
  up(s.instance_colors, tex.instance_colors.data(), tex.instance_colors.size() * 4);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (e != hipSuccess) {
    error = std::string("scene upload: ") + hipGetErrorString(e);
    return false;
  }
  s.vertex_words = sh.vertex_words;
  s.index_words = sh.index_words;
  s.material_count = M;
  const BuildInput in{s.positions, s.triangles, s.geometry_vertex_offset, s.geometry_triangle_offset,
                      s.instance_geometry, s.instance_transform, s.instance_base,
                      s.covered_instances ? s.instance_covered : nullptr, I, sh.tri_count, sh.leaf_size,
                      sh.builder};
  if (!build_bvh(in, s, r->stream, build_ms, error)) return false;
  e = ensure_launch(r);
  if (e != hipSuccess) {
    error = std::string("spill area: ") + hipGetErrorString(e);
    return false;
  }
  set_view(r, sh, p->environment);
  r->light_table = std::move(table);
  r->emissive_triangles = static_cast<uint64_t>(
      std::count_if(r->light_table.lights.begin(), r->light_table.lights.end(), [](const LightRecord& l) {
        return l.instance != kEnvironmentLight && (l.instance & kAnalyticLight) == 0;
      }));
  r->has_scene = true;
  ++r->bvh_builds;
  ++r->material_writes;
  ++r->light_writes;
  r->uploaded_bytes += bytes;
  return true;
}

// The same topology: only changed sections are uploaded (moved geometries'
// positions, transforms, tables), then a refit when anything moved.
bool apply_changes(lotus_hip_renderer* r, const lotus_hip_scene_plan* p, const PlanShape& sh, const Changes& c,
                   double& upload_ms, double& build_ms, std::string& error) {
  const auto t0 = std::chrono::steady_clock::now();
  SceneBuffers& s = r->scene;
  const uint32_t I = sh.instances, M = sh.materials;
  LightTableHost table;
  build_light_table(p, table);
  const bool lights_changed = !(table == r->light_table);
  hipError_t e = hipSuccess;
  uint64_t bytes = 0;
  auto up = [&](void* dst, const void* src, size_t n) {
    if (e == hipSuccess && n) {
      e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, r->stream);
      bytes += n;
    }
  };
  for (uint32_t g : c.moved) {
    const uint64_t first = 3ull * p->geometry_vertex_offset[g];
    up(s.positions + first, p->positions + first, 12ull * p->geometry_vertex_count[g]);
  }
  if (c.transforms) up(s.instance_transform, p->instance_transform, 64ull * I);
  if (c.instance_materials) up(s.instance_material, p->instance_material, 4ull * I);
  // the instances whose coverage is tested, when their materials changed
  bool covered_changed = false;
  if (c.materials || c.instance_materials) {
    std::vector<uint32_t> covered = covered_flags(p);
    covered_changed = covered != s.covered_host;
    if (covered_changed) {
      s.covered_host = std::move(covered);
      s.covered_instances = static_cast<uint32_t>(std::count(s.covered_host.begin(), s.covered_host.end(), 1u));
      up(s.instance_covered, s.covered_host.data(), 4ull * I);
    }
  }
  if (c.materials) {
    if (s.material_count != M) {
      dev_free(s.materials);
      e = dev_alloc(s.materials, 36ull * M);
      s.material_count = M;
    }
    up(s.materials, p->materials, 144ull * M);
    s.gltf_materials = count_gltf_materials(p);
  }
  if (lights_changed && e == hipSuccess) e = upload_light_table(s, table, up);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (e != hipSuccess) {
    error = std::string("scene upload: ") + hipGetErrorString(e);
    return false;
  }
  if (c.refit) {
    const BuildInput in{s.positions, s.triangles, s.geometry_vertex_offset, s.geometry_triangle_offset,
                        s.instance_geometry, s.instance_transform, s.instance_base,
                        s.covered_instances ? s.instance_covered : nullptr, I, sh.tri_count, sh.leaf_size,
                        sh.builder};
    if (!refit_bvh(in, s, r->stream, build_ms, error)) return false;
    ++r->bvh_refits;
  } else if (covered_changed) {
    e = mark_covered(s, sh.tri_count, s.covered_instances ? s.instance_covered : nullptr, r->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) {
      error = std::string("coverage marks: ") + hipGetErrorString(e);
      return false;
    }
  }
  if (c.materials) ++r->material_writes;
  if (lights_changed) ++r->light_writes;
  r->uploaded_bytes += bytes;
  r->light_table = std::move(table);
  r->emissive_triangles = static_cast<uint64_t>(
      std::count_if(r->light_table.lights.begin(), r->light_table.lights.end(), [](const LightRecord& l) {
        return l.instance != kEnvironmentLight && (l.instance & kAnalyticLight) == 0;
      }));
  set_view(r, sh, p->environment);
  return true;
}

}  // namespace

namespace {

// lotus_hip_svgf_static's scratch: SVGF's filter buffers (the a-trous
// iterations' two halves, the per-pixel guides and unpacked normals, 64 B
// per pixel) for at most kStaticSvgfPixels pixels, 128 MiB in all. The array lives in this
// library's code object: each GPU the library runs on holds its own copy
// until the process ends. Host code never passes its name to a kernel: the
// name is a host-side stand-in, and a kernel given it would fault. Each
// device's address comes from hipGetSymbolAddress, called once per device
// with that device current, and only that address is used.
constexpr uint32_t kStaticSvgfPixels = 1u << 21;  // 2,097,152: 1920 x 1080 fits
static __device__ __align__(256) float4 scratch[4ull * kStaticSvgfPixels];

std::mutex g_scratch_mutex;
std::map<int, std::pair<float4*, std::string>> g_scratch;  // device -> (address, error)

// The current device's copy of `scratch`, resolved on the device's first call.
void resolve_scratch(int device, float4*& address, std::string& error) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch.find(device);
  if (it == g_scratch.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    std::string message;
    if (e != hipSuccess || p == nullptr)
      message = std::string("the static scratch buffer has no device address: ") + hipGetErrorString(e);
    it = g_scratch.emplace(device, std::make_pair(e == hipSuccess ? static_cast<float4*>(p) : nullptr, message))
             .first;
  }
  address = it->second.first;
  error = it->second.second;
}

}  // namespace

extern "C" {

int lotus_hip_probe(char* detail, size_t detail_size) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    set_detail(detail, detail_size, "no HIP device");
    return LOTUS_HIP_SKIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
    set_detail(detail, detail_size, "hipGetDeviceProperties failed");
    return LOTUS_HIP_SKIP;
  }
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_detail(detail, detail_size, std::string("device is ") + prop.gcnArchName + ", the code objects are gfx950");
    return LOTUS_HIP_SKIP;
  }
  set_detail(detail, detail_size, std::string(prop.name) + " (" + prop.gcnArchName + ")");
  return LOTUS_HIP_PASS;
}

lotus_hip_renderer* lotus_hip_create(int device, int* status, char* detail, size_t detail_size) {
  char probe[256];
  const int p = lotus_hip_probe(probe, sizeof probe);
  if (p != LOTUS_HIP_PASS) {
    if (status) *status = p;
    set_detail(detail, detail_size, probe);
    return nullptr;
  }
  int count = 0;
  (void)hipGetDeviceCount(&count);
  if (device < 0 || device >= count) {
    if (status) *status = LOTUS_HIP_FAIL;
    set_detail(detail, detail_size, "device " + std::to_string(device) + " does not exist");
    return nullptr;
  }
  auto* r = new lotus_hip_renderer;
  r->device = device;
  hipError_t e = hipSetDevice(device);
  hipDeviceProp_t prop;
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  if (e == hipSuccess) {
    r->cus = static_cast<uint32_t>(prop.multiProcessorCount);
    r->grid = grid_shape(r->cus);
    r->events.resize(1024);
    for (auto& ev : r->events)
      if (e == hipSuccess) e = hipEventCreate(&ev);
  }
  if (e == hipSuccess) e = dev_alloc(r->counters, frame_counter_words(LOTUS_HIP_MAX_BOUNCES + 4));
  if (e == hipSuccess) e = dev_alloc(r->albedo_tables, sizeof kAlbedoTables / sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(r->albedo_tables, kAlbedoTables, sizeof kAlbedoTables, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = dev_alloc(r->stats, 8);
  if (e == hipSuccess) resolve_scratch(device, r->static_scratch, r->static_scratch_error);
  if (e != hipSuccess) {
    if (status) *status = LOTUS_HIP_FAIL;
    set_detail(detail, detail_size, std::string("device setup: ") + hipGetErrorString(e));
    lotus_hip_destroy(r);
    return nullptr;
  }
  if (status) *status = LOTUS_HIP_PASS;
  set_detail(detail, detail_size, probe);
  return r;
}

void lotus_hip_destroy(lotus_hip_renderer* r) {
  if (!r) return;
  (void)hipSetDevice(r->device);
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  free_target(r);
  free_scene(r);
  cache_free(r->cache);
  free_temporal(r);
  dev_free(r->temporal_ids);
  dev_free(r->albedo_tables);
  dev_free(r->svgf_filter);
  dev_free(r->counters);
  dev_free(r->stats);
  dev_free(r->spill);
  for (auto& ev : r->events)
    if (ev) (void)hipEventDestroy(ev);
  if (r->join) (void)hipEventDestroy(r->join);
  if (r->batch_exec) (void)hipGraphExecDestroy(r->batch_exec);
  dev_free(r->batch_vars);
  dev_free(r->fused_table);
  dev_free(r->fused_queue);
  dev_free(r->fused_spill);
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
}

// An 8-bit environment image's texels decoded to RGBA32F as TexelRgb decodes
// them (gpu_scene.cpp:612-629): channel / 255, then for sRGB the colour
// channels' transfer function, in double, rounded once to float. Alpha is
// not read. Returns a refusal, or "" with `out` describing the decoded copy.
std::string decode_environment(const lotus_hip_environment_map& e, lotus_hip_environment_map& out,
                               std::vector<float>& texels) {
  if (e.format != LOTUS_HIP_TEXTURE_RGBA8_SRGB && e.format != LOTUS_HIP_TEXTURE_RGBA8_UNORM)
    return "unknown environment texel format";
  if (e.width == 0 || e.height == 0 || e.width > 16384 || e.height > 16384 || !e.texels)
    return "the environment image must be 1 to 16384 texels on a side";
  const size_t n = size_t(e.width) * e.height;
  const uint8_t* b = static_cast<const uint8_t*>(e.texels);
  float lut[256];
  for (int v = 0; v < 256; ++v) {
    const double x = v / 255.0;
    lut[v] = static_cast<float>(e.format == LOTUS_HIP_TEXTURE_RGBA8_SRGB
                                    ? (x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4))
                                    : x);
  }
  texels.resize(4 * n);
  for (size_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) texels[4 * i + c] = lut[b[4 * i + c]];
    texels[4 * i + 3] = static_cast<float>(b[4 * i + 3] / 255.0);
  }
  out = e;
  out.texels = texels.data();
  out.format = LOTUS_HIP_TEXTURE_RGBA32F;
  return "";
}

int lotus_hip_scene_apply(lotus_hip_renderer* r, const lotus_hip_scene_plan* p, lotus_hip_scene_evidence* ev,
                          char* detail, size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  // an 8-bit environment image enters as its decoded RGBA32F texels
  lotus_hip_scene_plan decoded_plan;
  lotus_hip_environment_map decoded_map;
  std::vector<float> decoded_texels;
  if (p && p->environment_map && p->environment_map->format != LOTUS_HIP_TEXTURE_RGBA32F) {
    const std::string bad = decode_environment(*p->environment_map, decoded_map, decoded_texels);
    if (!bad.empty()) {
      set_detail(detail, detail_size, bad);
      return LOTUS_HIP_FAIL;
    }
    decoded_plan = *p;
    decoded_plan.environment_map = &decoded_map;
    p = &decoded_plan;
  }
  const std::string invalid = validate_plan(p, r->has_scene ? p->unchanged : 0u);
  if (!invalid.empty()) {
    set_detail(detail, detail_size, invalid);
    return LOTUS_HIP_FAIL;
  }
  (void)hipSetDevice(r->device);
  const PlanShape shape = plan_shape(p);
  const Changes c = compare_plan(r, p, shape);
  if (c.update == LOTUS_HIP_SCENE_BUILD && r->has_scene && p->unchanged != 0) {
    // a plan that rebuilds is read whole: it is validated whole
    const std::string whole = validate_plan(p);
    if (!whole.empty()) {
      set_detail(detail, detail_size, whole);
      return LOTUS_HIP_FAIL;
    }
  }
  if (c.update == LOTUS_HIP_SCENE_UNCHANGED) {
    fill_evidence(r, ev, c.update, 0, 0.0, 0.0);
    set_detail(detail, detail_size, "");
    return LOTUS_HIP_PASS;
  }
  std::string error;
  double build_ms = 0.0, upload_ms = 0.0;
  const bool ok = c.update == LOTUS_HIP_SCENE_BUILD ? apply_build(r, p, shape, upload_ms, build_ms, error)
                                                     : apply_changes(r, p, shape, c, upload_ms, build_ms, error);
  if (!ok) {
    // A partly applied plan leaves the scene unknown (vulkan_backend.cpp:2629-2633).
    free_scene(r);
    r->failure = error;
    set_detail(detail, detail_size, error);
    return LOTUS_HIP_FAIL;
  }
  record_plan(r, p, shape, c);
  // Any change restarts the accumulation (vulkan_backend.cpp:2621-2626);
  // temporal history survives only a compatible one (:262-289, 1782-1788).
  ++r->scene_generation;
  if (!c.temporal_compatible) ++r->scene_epoch;
  r->rounds_hint = 0;
  if (c.update != LOTUS_HIP_SCENE_TABLES) r->bounds_valid = false;
  fill_evidence(r, ev, c.update, static_cast<uint32_t>(c.moved.size()), upload_ms, build_ms);
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

namespace {
int render_impl(lotus_hip_renderer* r, const lotus_hip_frame_params* P, lotus_hip_frame_result* res, char* detail,
                size_t detail_size, TemporalFrameState* tf, bool deferred = false);
}  // namespace

int lotus_hip_render(lotus_hip_renderer* r, const lotus_hip_frame_params* P, lotus_hip_frame_result* res,
                     char* detail, size_t detail_size) {
  return render_impl(r, P, res, detail, detail_size, nullptr);
}

}  // extern "C"

namespace {
// DisplayViewport and DataScissor (vulkan_backend.cpp:651-677): a display
// window of positive size, else the whole image; the data window clipped to
// the image (an empty one renders nothing), or the whole image when its size
// is not positive.
void resolved_windows(const lotus_hip_frame_params* P, float* display, uint32_t* data) {
  const float* dw = P->display_window;
  const bool whole = dw[2] <= 0.0f || dw[3] <= 0.0f;
  display[0] = whole ? 0.0f : dw[0];
  display[1] = whole ? 0.0f : dw[1];
  display[2] = whole ? static_cast<float>(P->width) : dw[2];
  display[3] = whole ? static_cast<float>(P->height) : dw[3];
  data[0] = 0;
  data[1] = 0;
  data[2] = P->width;
  data[3] = P->height;
  const int32_t* w = P->data_window;
  if (w[2] > 0 && w[3] > 0) {
    const int64_t x0 = std::max<int64_t>(w[0], 0), y0 = std::max<int64_t>(w[1], 0);
    const int64_t x1 = std::min<int64_t>(int64_t{w[0]} + w[2], P->width);
    const int64_t y1 = std::min<int64_t>(int64_t{w[1]} + w[3], P->height);
    if (x1 <= x0 || y1 <= y0) {
      data[0] = data[1] = data[2] = data[3] = 0;
    } else {
      data[0] = static_cast<uint32_t>(x0);
      data[1] = static_cast<uint32_t>(y0);
      data[2] = static_cast<uint32_t>(x1 - x0);
      data[3] = static_cast<uint32_t>(y1 - y0);
    }
  }
}

// AccumulationKey and restart (vulkan_backend.cpp:419-434, 1186-1205): the
// accumulation starts over when anything it depends on changed.
void restart_if_changed(lotus_hip_renderer* r, const lotus_hip_frame_params* P, const float* vk_world_to_clip,
                        bool cache_frame, bool restir) {
  AccumulationKey key;
  std::memcpy(key.world_to_clip.data(), vk_world_to_clip, 64);
  key.width = P->width;
  key.height = P->height;
  std::memcpy(key.display.data(), P->display_window, 16);
  std::memcpy(key.data.data(), P->data_window, 16);
  key.light_sampling = P->light_sampling;
  key.sample_index = P->sample_index;
  key.max_bounces = P->max_bounces;
  key.scene_generation = r->scene_generation;
  key.cache_frames = cache_frame ? 1u : 0u;
  key.restir_candidates = restir ? P->restir_candidates : 0u;
  key.restir_proposal = restir ? P->restir_proposal : 0u;
  if (!r->key_valid || !(r->key == key)) {
    r->key = key;
    r->key_valid = true;
    r->accumulated = 0;
  }
}

// `deferred` (batched frames): one sample is enqueued on the renderer's
// stream with the previous sample's rounds, without events, readbacks or a
// synchronisation; batch_finish reads its counters afterwards.
int render_impl(lotus_hip_renderer* r, const lotus_hip_frame_params* P, lotus_hip_frame_result* res, char* detail,
                size_t detail_size, TemporalFrameState* tf, bool deferred) {
  if (!r || !P) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  // Settings errors fail without latching (vulkan_backend.cpp:1034-1090).
  if (P->output > LOTUS_HIP_OUTPUT_ROUGHNESS_METALLIC) {
    set_detail(detail, detail_size, "unknown scene output");
    return LOTUS_HIP_FAIL;
  }
  if (P->output == LOTUS_HIP_OUTPUT_RADIANCE && P->max_bounces > LOTUS_HIP_MAX_BOUNCES) {
    set_detail(detail, detail_size, "the wavefront integrator traces at most 1024 bounces");
    return LOTUS_HIP_FAIL;
  }
  if (P->light_sampling > LOTUS_HIP_LIGHT_RESTIR) {
    set_detail(detail, detail_size, "unknown light sampling");
    return LOTUS_HIP_FAIL;
  }
  if (P->light_sampling == LOTUS_HIP_LIGHT_RESTIR && (P->restir_candidates == 0 || P->restir_candidates > 64)) {
    set_detail(detail, detail_size, "initial reservoir candidates must be in [1, 64]");
    return LOTUS_HIP_FAIL;
  }
  if (P->restir_proposal > 2) {
    set_detail(detail, detail_size, "unknown reservoir proposal");
    return LOTUS_HIP_FAIL;
  }
  if (P->width == 0 || P->height == 0) {
    set_detail(detail, detail_size, "the target is empty");
    return LOTUS_HIP_FAIL;
  }
  if (!r->has_scene) {
    set_detail(detail, detail_size, "no scene has been applied");
    return LOTUS_HIP_FAIL;
  }
  const bool cache_frame = (P->flags & LOTUS_HIP_FRAME_RADIANCE_CACHE) != 0 && P->output == LOTUS_HIP_OUTPUT_RADIANCE;
  if (cache_frame && !r->cache.ready) {
    set_detail(detail, detail_size, "radiance cache frames need lotus_hip_cache_configure first");
    return LOTUS_HIP_FAIL;
  }
  FrameConst F{};
  vulkan_world_to_clip(P->world_to_clip, F.world_to_clip);
  if (!invert(F.world_to_clip, F.clip_to_world)) {
    set_detail(detail, detail_size, "ray-query camera is singular or non-finite");
    return LOTUS_HIP_FAIL;
  }
  (void)hipSetDevice(r->device);
  // DisplayViewport / DataScissor, and ndc_from_pixel (vulkan_backend.cpp:2408-2411)
  float viewport[4];
  uint32_t rect[4];
  resolved_windows(P, viewport, rect);
  const float vx = viewport[0], vy = viewport[1], vw = viewport[2], vh = viewport[3];
  F.ndc_from_pixel[0] = 2.0f / vw;
  F.ndc_from_pixel[1] = 2.0f / vh;
  F.ndc_from_pixel[2] = -2.0f * vx / vw - 1.0f;
  F.ndc_from_pixel[3] = -2.0f * vy / vh - 1.0f;
  std::memcpy(F.rect, rect, sizeof rect);
  std::memcpy(F.background, P->clear_color, 16);
  F.width = P->width;
  F.height = P->height;
  F.max_bounces = P->max_bounces;
  F.light_sampling = P->light_sampling;
  F.output = P->output;

  hipError_t e = ensure_target(r, P->width, P->height);
  if (e == hipSuccess && cache_frame) e = ensure_cache_target(r);
  const bool restir = P->light_sampling == LOTUS_HIP_LIGHT_RESTIR && P->output == LOTUS_HIP_OUTPUT_RADIANCE;
  if (e == hipSuccess && restir && !r->q.restir) e = dev_alloc(r->q.restir, 4ull * P->width * P->height);
  if (e == hipSuccess && cache_frame) e = ensure_bounds(r);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "target allocation", e));
    return LOTUS_HIP_FAIL;
  }
  if (cache_frame) {
    F.cache_depth = r->cache_config.query_depth;
    F.train_stride = r->train_stride;
    F.train_capacity = r->train_capacity;
    for (int a = 0; a < 3; ++a) {
      F.cache_lo[a] = r->scene_lo[a];
      F.cache_scale[a] = r->scene_scale[a];
    }
  }
  F.restir_candidates = P->restir_candidates;
  F.restir_proposal = P->restir_proposal;
  F.restir = r->q.restir;
  const float4* final_records = r->q.restir;
  if (tf) {
    std::memcpy(F.view, tf->view, 64);
    F.restir_surf = tf->reuse ? const_cast<float4*>(tf->A.surf) : nullptr;
    F.instance_ids = tf->I.instance_id;
    F.material_ids = tf->I.material_id;
    // a fresh batch never continues a progressive mean; a settling frame
    // continues the still frames' progressive accumulation, and a frame
    // without accumulation or reuse the ordinary one
    if (tf->fresh && !tf->settling) r->key_valid = false;
  }
  uint64_t cache_queries = 0, cache_samples = 0, cache_used = 0;
  uint32_t last_training = 0;  // the last sample's training records
  double cache_loss_sum = 0.0;
  const uint32_t npix = P->width * P->height;
  Launch L = launch_for(r, r->stream);
  const bool split = (P->flags & LOTUS_HIP_FRAME_SPLIT_KINDS) != 0;
  launch_zero(reinterpret_cast<uint32_t*>(r->stats), 16, r->stream);
  launch_fill4(r->q.color, npix, make_float4(P->clear_color[0], P->clear_color[1], P->clear_color[2], P->clear_color[3]),
               r->stream);
  if (res && res->depth) launch_fill1(r->depth, npix, 1.0f, r->stream);

  double t_ms[kTimers] = {};
  double frame_ms = 0.0;
  if (tf) {
    // the pixel-centre surfaces and their reprojection, before any pass reads them
    hipEvent_t* ev = r->events.data() + r->events.size() - 4;
    (void)hipEventRecord(ev[0], r->stream);
    launch_fill4(reinterpret_cast<float4*>(tf->surface), 4u * npix, make_float4(0.0f, 0.0f, 0.0f, 0.0f), r->stream);
    launch_temporal_surface(L, r->view, F, tf->T, tf->I, tf->surface);
    launch_temporal_reproject(L, tf->T, tf->surface, tf->previous_surface, tf->validity);
    (void)hipEventRecord(ev[1], r->stream);
  }
  uint32_t added = 0;
  std::vector<uint32_t> counts;
  uint64_t camera_rays = 0, extension_rays = 0, shadow_rays = 0;
  uint64_t round_extension_rays = 0, round_shadow_rays = 0;  // walked by the rounds' traversal launches
  uint64_t restir_rays = 0;
  uint64_t trace_launches = 0;
  uint32_t rounds_used = 0;
  if (P->output == LOTUS_HIP_OUTPUT_RADIANCE) {
    restart_if_changed(r, P, F.world_to_clip, cache_frame, restir);
    const uint32_t room = P->max_samples == 0 ? P->frame_count
                                              : P->max_samples - std::min(P->max_samples, r->accumulated);
    const uint32_t to_add = std::min(P->frame_count, room);
    const uint32_t submissions = std::max(to_add, 1u);
    const uint32_t bounces = P->max_bounces + 1;
    const uint32_t stride = P->max_bounces + 2;
    for (uint32_t sample = 0; sample < submissions && e == hipSuccess; ++sample) {
      const uint32_t add = sample < to_add ? 1u : 0u;
      F.sample_index = P->sample_index + r->accumulated;
      F.accumulated = r->accumulated;
      if (deferred && r->graph_batch && r->graph_vars) {
        // a captured sample reads its frame variables from device memory,
        // so a replay of the graph takes another frame's values
        F.vars = r->graph_vars;
        r->graph_host_vars[0] = F.sample_index;
        r->graph_host_vars[1] = F.accumulated;
        r->graph_host_vars[2] = add;
        r->graph_host_vars[3] = 0u;
      }
      uint32_t rounds = r->rounds_hint == 0 ? std::min(bounces, kFirstRounds) : std::min(r->rounds_hint, bounces);
      // a graphed batch runs every bounce as a round: its launches, and so
      // its graph's topology, stay the same from frame to frame (results do
      // not depend on the rounds)
      if (deferred && r->graph_batch) rounds = std::min(bounces, kFirstRounds);
      rounds = std::min(std::max(rounds, 1u), kMaxRounds);
      if (cache_frame) {
        rounds = std::min(std::max(rounds, F.cache_depth + 1), std::min(bounces, kMaxRounds));
        F.train_offset = mix_host(F.sample_index * 0x9e3779b9u + 0x632be5abu) % r->train_stride;
      }
      size_t ev = 0;
      std::vector<std::pair<int, size_t>> pairs;
      auto mark = [&](int timer) {
        if (deferred) return;
        pairs.emplace_back(timer, ev);
        (void)hipEventRecord(r->events[ev++], r->stream);
      };
      auto done = [&]() {
        if (!deferred) (void)hipEventRecord(r->events[ev++], r->stream);
      };
      launch_zero(r->counters, frame_counter_words(stride), r->stream);
      if (cache_frame) {
        launch_zero(r->q.crec_n, r->train_capacity, r->stream);
        launch_fill1(r->cache.loss, 1, 0.0f, r->stream);
      }
      // the path megakernel: each path from its camera ray to its end in one
      // launch (its timer is the tail's)
      const bool mega = add && r->tuning.megakernel != 0 && !split && !tf && paths_eligible(r->view, F);
      if (mega && r->fuse && deferred) {
        // a fused batch launches the paths of all its views itself, then
        // each view's resolve
        r->fused_view = PathView{r->view, F, r->q, r->counters, r->stats, 0u, path_count(F)};
        r->fused_add = add;
        r->fused_pending = true;
        r->deferred = {true, add != 0, stride, 0u, 0u};
        return LOTUS_HIP_PASS;
      }
      if (mega) {
        mark(kTTail);
        launch_paths(L, r->view, F, r->q, r->counters);
        done();
        rounds = 0;
      } else if (add) {
        if (tf && tf->reuse) launch_restir_clear(L, r->q.restir, const_cast<float4*>(tf->A.surf), npix);
        mark(kTCamera);
        launch_camera(L, r->view, F, r->q, r->counters);
        done();
        if (restir) {
          mark(kTRestir);
          F.restir = r->q.restir;
          if (tf) tf->A.initial = r->q.restir;
          launch_restir_initial(L, r->view, F, r->q, r->counters);
          if (tf && tf->reuse) {
            launch_restir_temporal(L, r->view, F, tf->A, r->counters);
            if (tf->spatial) launch_restir_spatial(L, r->view, F, tf->A, r->counters);
            F.restir = tf->spatial ? tf->A.spatial_out : tf->A.temporal_out;
            final_records = F.restir;
          }
          done();
        }
        // round b: shade bounce b, then trace the extension rays of bounce
        // b + 1 (unless the tail takes them) with the shadow rays of bounce b.
        // Before a tail, the last round's shadow rays go to the tail too: a
        // path's pair walks it beside the path's next extension ray, and the
        // tail walks the shadow rays of paths that end there on their own.
        const bool tail = rounds < bounces;
        for (uint32_t b = 0; b < rounds; ++b) {
          const bool link = tail && b + 1 == rounds;
          mark(kTShade);
          launch_shade(L, r->view, F, r->q, r->counters, b, link);
          done();
          // shadow rays: next event estimation, or under BSDF sampling a
          // parallel light's
          const bool shadows = P->light_sampling != 0 || r->view.parallel != 0;
          const uint32_t kinds = (b + 1 < rounds ? kTraceExtension : 0u) | (shadows && !link ? kTraceShadow : 0u);
          if (kinds == 0) continue;
          const uint32_t weigh = P->light_sampling != 0 ? kTraceWeigh : 0u;
          if (split) {
            for (const uint32_t kind : {kTraceExtension, kTraceShadow}) {
              if (!(kinds & kind)) continue;
              mark(kind == kTraceExtension ? kTIntersect : kTShadow);
              launch_trace(L, r->view, r->q, r->counters, b + 1, kind | weigh);
              if (kind == kTraceExtension && scene_has_extras(r->view))
                launch_escape(L, r->view, r->q, r->counters, b + 1, weigh != 0);
              done();
              if (add) ++trace_launches;
            }
          } else {
            mark(kTTrace);
            launch_trace(L, r->view, r->q, r->counters, b + 1, kinds | weigh);
            if ((kinds & kTraceExtension) && scene_has_extras(r->view))
              launch_escape(L, r->view, r->q, r->counters, b + 1, weigh != 0);
            done();
            if (add) ++trace_launches;
          }
        }
        if (tail) {
          mark(kTTail);
          launch_tail(L, r->view, F, r->q, r->counters, rounds);
          done();
        }
        if (cache_frame) {
          mark(kTCacheInfer);
          cache_infer_add(r->cache, r->q.cq, r->q.cq_thr, cache_query_count(r->counters), r->q.rad, r->stream);
          done();
        }
      }
      mark(kTResolve);
      launch_resolve(L, F, r->q, add);
      done();
      if (cache_frame && add) {
        // train on this frame's paths: the samples, then train_steps Adam
        // steps over interleaved subsets of them
        const uint32_t steps = std::max(1u, r->cache_config.train_steps);
        mark(kTCacheTrain);
        launch_cache_samples(L, F, r->q, r->counters, r->sample_capacity);
        for (uint32_t k = 0; k < steps; ++k) {
          cache_backprop(r->cache, r->q.samples, r->sample_capacity, cache_sample_count(r->counters), k, steps,
                         r->stream);
          cache_adam(r->cache, r->sample_capacity, cache_sample_count(r->counters), k, steps, r->stream);
        }
        done();
      }
      float loss_host = 0.0f;
      if (cache_frame && add) {
        e = hipGetLastError();
        if (e == hipSuccess)
          e = hipMemcpyAsync(&loss_host, r->cache.loss, sizeof(float), hipMemcpyDeviceToHost, r->stream);
      }
      if (deferred) {
        r->deferred = {true, add != 0, stride, rounds, trace_launches};
        return LOTUS_HIP_PASS;
      }
      std::vector<uint32_t> host(frame_counter_words(stride));
      e = hipGetLastError();
      if (e == hipSuccess)
        e = hipMemcpyAsync(host.data(), r->counters, host.size() * 4, hipMemcpyDeviceToHost, r->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
      if (e != hipSuccess) break;
      for (const auto& [timer, first] : pairs) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, r->events[first], r->events[first + 1]);
        t_ms[timer] += ms;
      }
      if (ev >= 2) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, r->events[0], r->events[ev - 1]);
        frame_ms += ms;
      }
      if (cache_frame && add) {
        const uint32_t total = std::min(host[counter_word(1, 3)], r->sample_capacity);
        last_training = total;
        const uint32_t steps = std::max(1u, r->cache_config.train_steps);
        for (uint32_t k = 0; k < steps; ++k)
          if (total > k) cache_used += std::min((total - k + steps - 1) / steps, r->cache_config.batch_capacity);
        cache_queries += host[counter_word(0, 3)];
        cache_samples += host[counter_word(1, 3)];
        cache_loss_sum += loss_host;
      }
      if (add) {
        counts.assign(stride, 0u);
        size_t last = 0;
        for (uint32_t b = 0; b < stride; ++b) {
          counts[b] = host[counter_word(b, 0)];
          if (counts[b]) last = b;
          if (b >= 1) extension_rays += host[counter_word(b, 0)];
          shadow_rays += host[counter_word(b, 2)];
          if (b >= 1 && b < rounds) round_extension_rays += host[counter_word(b, 0)];
          if (b + 1 < rounds || (b < rounds && rounds >= bounces)) round_shadow_rays += host[counter_word(b, 2)];
        }
        camera_rays += host[0];
        shadow_rays += *restir_ray_count(host.data()) + *parallel_ray_count(host.data());
        restir_rays += *restir_ray_count(host.data());
        counts.resize(host[0] ? last + 1 : 0);
        r->rounds_hint = wavefront_rounds(counts, r->tuning.tail_divisor);
        rounds_used = rounds;
        ++r->accumulated;
        ++added;
      }
    }
    if (e != hipSuccess) {
      set_detail(detail, detail_size, latch(r, "wavefront frame", e));
      return LOTUS_HIP_FAIL;
    }
  }
  if (tf && tf->batch_host && e == hipSuccess)
    e = hipMemcpyAsync(tf->batch_host, r->q.color, npix * 16ull, hipMemcpyDeviceToHost, r->stream);
  if (tf && tf->accumulate) {
    // SVGF over this call's batch (the resolved colour): accumulation,
    // variance, a-trous iterations and the composite back into the colour;
    // a settling frame crossfades its progressive mean from the prior instead
    hipEvent_t* ev = r->events.data() + r->events.size() - 4;
    (void)hipEventRecord(ev[2], r->stream);
    if (tf->settling) {
      launch_temporal_settle(L, tf->T, tf->prior, tf->old_radiance, r->q.color, r->accumulated, tf->max_samples);
    } else {
      launch_temporal_accumulate(L, tf->T, r->q.color, tf->surface, tf->validity, tf->old_radiance, tf->radiance);
      launch_temporal_variance(L, tf->T, tf->surface, tf->radiance, tf->filter);
      launch_temporal_filter(L, tf->T, tf->surface, tf->radiance, tf->filter, r->q.color);
    }
    (void)hipEventRecord(ev[3], r->stream);
  }
  // The pixel-centre camera pass: depth, and the colour of the other outputs.
  const bool diag = P->output != LOTUS_HIP_OUTPUT_RADIANCE;
  if (diag || (res && res->depth)) {
    F.sample_index = 0;
    launch_camera_pass(L, r->view, F, r->q, (res && res->depth) ? r->depth : nullptr, diag ? 1u : 0u);
  }
  if (res) {
    if (e == hipSuccess && res->color)
      e = hipMemcpyAsync(res->color, r->q.color, npix * 16ull, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess && res->depth)
      e = hipMemcpyAsync(res->depth, r->depth, npix * 4ull, hipMemcpyDeviceToHost, r->stream);
    if (e == hipSuccess && res->restir_records && restir && final_records)
      e = hipMemcpy2DAsync(res->restir_records, 48, final_records, 64, 48, npix, hipMemcpyDeviceToHost, r->stream);
  }
  unsigned long long stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(stats, r->stats, sizeof stats, hipMemcpyDeviceToHost, r->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "frame readback", e));
    return LOTUS_HIP_FAIL;
  }
  if (res) {
    res->path_counts_size = static_cast<uint32_t>(std::min<size_t>(counts.size(), res->path_counts_capacity));
    if (res->path_counts)
      std::memcpy(res->path_counts, counts.data(), res->path_counts_size * 4ull);
    res->wavefront_rounds = counts.empty() ? 0 : rounds_used;
    res->samples_per_pixel = P->output == LOTUS_HIP_OUTPUT_RADIANCE ? r->accumulated : 0;
    res->frames_rendered = std::max(added, 1u);
    res->camera_rays = camera_rays;
    res->extension_rays = extension_rays;
    res->shadow_rays = shadow_rays;
    res->camera_gpu_ms = t_ms[kTCamera];
    res->intersect_gpu_ms = t_ms[kTIntersect];
    res->trace_gpu_ms = t_ms[kTTrace];
    res->tail_gpu_ms = t_ms[kTTail];
    res->shade_gpu_ms = t_ms[kTShade];
    res->shadow_gpu_ms = t_ms[kTShadow];
    res->resolve_gpu_ms = t_ms[kTResolve];
    res->frame_gpu_ms = frame_ms;
    for (int k = 0; k < 4; ++k) {
      res->traversal_nodes[k] = stats[2 * k];
      res->traversal_tris[k] = stats[2 * k + 1];
    }
    res->trace_launches = trace_launches;
    res->traversal_rays[0] = camera_rays;
    res->traversal_rays[1] = round_extension_rays;
    res->traversal_rays[2] = round_shadow_rays;
    res->traversal_rays[3] =
        (extension_rays - round_extension_rays) + (shadow_rays - restir_rays - round_shadow_rays);
    res->cache_queries = cache_queries;
    res->cache_samples = cache_samples;
    res->cache_loss = cache_used ? cache_loss_sum / (3.0 * static_cast<double>(cache_used)) : 0.0;
    res->cache_infer_gpu_ms = t_ms[kTCacheInfer];
    res->cache_train_gpu_ms = t_ms[kTCacheTrain];
    res->restir_gpu_ms = t_ms[kTRestir];
    res->cache_training_count = 0;
    if (res->cache_training && cache_frame && last_training) {
      const uint32_t n = std::min(last_training, res->cache_training_capacity);
      const hipError_t ce = hipMemcpy(res->cache_training, r->q.samples, size_t(n) * kSampleFloats * 4,
                                      hipMemcpyDeviceToHost);
      if (ce != hipSuccess) {
        set_detail(detail, detail_size, latch(r, "training record readback", ce));
        return LOTUS_HIP_FAIL;
      }
      res->cache_training_count = n;
    }
    if (tf) {
      hipEvent_t* ev = r->events.data() + r->events.size() - 4;
      float a = 0.0f, b = 0.0f;
      (void)hipEventElapsedTime(&a, ev[0], ev[1]);
      if (tf->accumulate) (void)hipEventElapsedTime(&b, ev[2], ev[3]);
      res->frame_gpu_ms += a + b;
      res->temporal_surface_gpu_ms = a;
      res->svgf_gpu_ms = b;
    } else {
      res->temporal_surface_gpu_ms = 0.0;
      res->svgf_gpu_ms = 0.0;
    }
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}
}  // namespace

extern "C" {

int lotus_hip_trace_device(lotus_hip_renderer* r, uint32_t count, const void* rays, void* hits, int any_hit,
                           void* stream, char* detail, size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  if (!r->has_scene) {
    set_detail(detail, detail_size, "no scene has been applied");
    return LOTUS_HIP_FAIL;
  }
  if (count == 0) return LOTUS_HIP_PASS;
  if (!rays || !hits) {
    set_detail(detail, detail_size, "null ray or hit buffer");
    return LOTUS_HIP_FAIL;
  }
  (void)hipSetDevice(r->device);
  Launch L = launch_for(r, stream ? static_cast<hipStream_t>(stream) : r->stream);
  launch_rays(L, r->view, static_cast<const float*>(rays), static_cast<uint32_t*>(hits), count,
               any_hit == 2 ? 2u : any_hit ? 1u : 0u);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "trace launch", e));
    return LOTUS_HIP_FAIL;
  }
  return LOTUS_HIP_PASS;
}

int lotus_hip_trace(lotus_hip_renderer* r, uint32_t count, const float* rays, uint32_t* hits, int any_hit,
                    char* detail, size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (count == 0) return LOTUS_HIP_PASS;
  (void)hipSetDevice(r->device);
  float* d_rays = nullptr;
  uint32_t* d_hits = nullptr;
  const size_t out_words = any_hit ? count : 5ull * count;
  hipError_t e = dev_alloc(d_rays, 8ull * count);
  if (e == hipSuccess) e = dev_alloc(d_hits, out_words);
  if (e == hipSuccess) e = hipMemcpyAsync(d_rays, rays, 32ull * count, hipMemcpyHostToDevice, r->stream);
  if (e == hipSuccess) {
    const int s = lotus_hip_trace_device(r, count, d_rays, d_hits, any_hit, r->stream, detail, detail_size);
    if (s != LOTUS_HIP_PASS) {
      dev_free(d_rays);
      dev_free(d_hits);
      return s;
    }
    e = hipMemcpyAsync(hits, d_hits, out_words * 4, hipMemcpyDeviceToHost, r->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  dev_free(d_rays);
  dev_free(d_hits);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "trace", e));
    return LOTUS_HIP_FAIL;
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

int lotus_hip_set_tuning(lotus_hip_renderer* r, const lotus_hip_tuning* tuning, char* detail, size_t detail_size) {
  if (!r || !tuning) {
    set_detail(detail, detail_size, "no renderer or tuning");
    return LOTUS_HIP_FAIL;
  }
  if (tuning->resident_percent > 100) {
    set_detail(detail, detail_size, "resident_percent must be in [1, 100] (0 for 100)");
    return LOTUS_HIP_FAIL;
  }
  if (tuning->camera_tile > 2) {
    set_detail(detail, detail_size, "camera_tile must be 0 (8 x 8), 1 (16 x 4) or 2 (4 x 16)");
    return LOTUS_HIP_FAIL;
  }
  r->tuning = *tuning;
  r->rounds_hint = 0;  // the next sample sizes the rounds under the new rule
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

int lotus_hip_get_tuning(lotus_hip_renderer* r, lotus_hip_tuning* tuning) {
  if (!r || !tuning) return LOTUS_HIP_FAIL;
  *tuning = r->tuning;
  return LOTUS_HIP_PASS;
}

int lotus_hip_light_table(lotus_hip_renderer* r, uint32_t* offsets, uint32_t* records, uint32_t* count,
                          char* detail, size_t detail_size) {
  if (!r || !r->has_scene) {
    set_detail(detail, detail_size, "no scene has been applied");
    return LOTUS_HIP_FAIL;
  }
  // The tables the kernels read, copied from the device.
  const SceneBuffers& s = r->scene;
  if (count) *count = s.light_count;
  (void)hipSetDevice(r->device);
  hipError_t e = hipSuccess;
  if (offsets && r->view.instance_count)
    e = hipMemcpy(offsets, s.light_offsets, size_t(r->view.instance_count) * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && records && s.light_count)
    e = hipMemcpy(records, s.lights, size_t(s.light_count) * sizeof(LightRecord), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "light table readback", e));
    return LOTUS_HIP_FAIL;
  }
  return LOTUS_HIP_PASS;
}

int lotus_hip_readback_scene(lotus_hip_renderer* r, lotus_hip_scene_contents* out, char* detail,
                             size_t detail_size) {
  if (!r || !out) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  if (!r->has_scene) {
    set_detail(detail, detail_size, "no scene has been applied");
    return LOTUS_HIP_FAIL;
  }
  const SceneBuffers& s = r->scene;
  const auto& plan = r->plan;
  const uint32_t G = static_cast<uint32_t>(plan.gvo.size()), I = r->view.instance_count;
  out->geometry_count = G;
  out->instance_count = I;
  out->material_count = s.material_count;
  out->triangle_count = r->view.tri_count;
  out->node_count = s.node_count;
  out->light_count = s.light_count;
  out->position_floats = s.vertex_words;
  out->index_words = s.index_words;
  for (int c = 0; c < 3; ++c) out->environment[c] = r->view.environment[c];
  (void)hipSetDevice(r->device);
  hipError_t e = hipSuccess;
  auto down = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess && dst && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  };
  down(out->positions, s.positions, s.vertex_words * 4);
  down(out->triangles, s.triangles, s.index_words * 4);
  down(out->geometry_vertex_offset, s.geometry_vertex_offset, G * 4ull);
  down(out->geometry_triangle_offset, s.geometry_triangle_offset, G * 4ull);
  down(out->instance_geometry, s.instance_geometry, I * 4ull);
  down(out->instance_transform, s.instance_transform, I * 64ull);
  down(out->instance_material, s.instance_material, I * 4ull);
  down(out->materials, s.materials, s.material_count * 144ull);
  down(out->bvh_nodes, s.nodes, size_t(s.node_count) * sizeof(WideNode));
  down(out->bvh_triangles, s.tris, size_t(r->view.tri_count) * sizeof(Tri));
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "scene readback", e));
    return LOTUS_HIP_FAIL;
  }
  if (out->geometry_vertex_count && G) std::memcpy(out->geometry_vertex_count, plan.gvc.data(), G * 4ull);
  if (out->geometry_triangle_count && G) std::memcpy(out->geometry_triangle_count, plan.gtc.data(), G * 4ull);
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

int lotus_hip_read_accumulation(lotus_hip_renderer* r, float* sums, uint32_t* samples, char* detail,
                                size_t detail_size) {
  if (!r || !r->q.accum) {
    set_detail(detail, detail_size, "no frame has been rendered");
    return LOTUS_HIP_FAIL;
  }
  if (samples) *samples = r->accumulated;
  if (sums) {
    const hipError_t e = hipMemcpy(sums, r->q.accum, size_t(r->width) * r->height * 16, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      set_detail(detail, detail_size, latch(r, "accumulation readback", e));
      return LOTUS_HIP_FAIL;
    }
  }
  return LOTUS_HIP_PASS;
}

namespace {

bool cache_usable(lotus_hip_renderer* r, char* detail, size_t detail_size) {
  if (!r) return false;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return false;
  }
  if (!r->cache.ready) {
    set_detail(detail, detail_size, "the radiance cache is not configured");
    return false;
  }
  (void)hipSetDevice(r->device);
  return true;
}

int cache_finish(lotus_hip_renderer* r, hipError_t e, const char* what, char* detail, size_t detail_size) {
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, what, e));
    return LOTUS_HIP_FAIL;
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

}  // namespace

int lotus_hip_cache_configure(lotus_hip_renderer* r, lotus_hip_cache_config* config, char* detail,
                              size_t detail_size) {
  if (!r || !config) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  if (config->query_depth < 1 || config->train_paths < 1 || config->train_steps < 1 || config->train_steps > 64) {
    set_detail(detail, detail_size, "query_depth and train_paths must be at least 1, train_steps in [1, 64]");
    return LOTUS_HIP_FAIL;
  }
  CacheConfig c;
  c.log2_table = config->log2_table;
  c.min_resolution = config->min_resolution;
  c.max_resolution = config->max_resolution;
  c.batch_capacity = config->batch_capacity;
  c.learning_rate = config->learning_rate;
  c.beta1 = config->beta1;
  c.beta2 = config->beta2;
  c.epsilon = config->epsilon;
  c.seed = config->seed;
  (void)hipSetDevice(r->device);
  (void)hipStreamSynchronize(r->stream);
  std::string err;
  if (!cache_init(r->cache, c, r->stream, err)) {
    set_detail(detail, detail_size, err);
    return LOTUS_HIP_FAIL;
  }
  for (int l = 0; l < kCacheLevels; ++l) config->resolution[l] = r->cache.resolution[l];
  config->dense_levels = r->cache.dense_mask;
  r->cache_config = *config;
  r->key_valid = false;
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

int lotus_hip_cache_infer(lotus_hip_renderer* r, const void* queries_device, uint32_t count, void* radiance_device,
                          char* detail, size_t detail_size) {
  if (!cache_usable(r, detail, detail_size)) return LOTUS_HIP_FAIL;
  if (count == 0) return LOTUS_HIP_PASS;
  if (!queries_device || !radiance_device) {
    set_detail(detail, detail_size, "queries and radiance must be device buffers");
    return LOTUS_HIP_FAIL;
  }
  cache_infer(r->cache, static_cast<const float*>(queries_device), count, static_cast<float*>(radiance_device),
              r->stream);
  return cache_finish(r, hipSuccess, "cache inference", detail, detail_size);
}

int lotus_hip_cache_gradients(lotus_hip_renderer* r, const void* samples_device, uint32_t count, float* weight_grad,
                              float* grid_grad, float* loss, char* detail, size_t detail_size) {
  if (!cache_usable(r, detail, detail_size)) return LOTUS_HIP_FAIL;
  if (count == 0 || count > r->cache.config.batch_capacity || !samples_device) {
    set_detail(detail, detail_size, "count must lie in [1, batch_capacity] with a device sample buffer");
    return LOTUS_HIP_FAIL;
  }
  CacheState& c = r->cache;
  launch_fill1(c.loss, 1, 0.0f, r->stream);
  cache_backprop(c, static_cast<const float*>(samples_device), count, nullptr, 0, 1, r->stream);
  if (grid_grad) cache_merge_grid_grad(c, r->stream);
  std::vector<float> wg(kCacheParams);
  float loss_sum = 0.0f;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(wg.data(), c.wgrad, sizeof(float) * kCacheParams, hipMemcpyDeviceToHost, r->stream);
  if (e == hipSuccess && grid_grad)
    e = hipMemcpyAsync(grid_grad, c.ggrad, sizeof(float) * c.grid_entries, hipMemcpyDeviceToHost, r->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&loss_sum, c.loss, sizeof(float), hipMemcpyDeviceToHost, r->stream);
  if (e == hipSuccess) cache_clear_grid_grad(c, r->stream);
  const int status = cache_finish(r, e, "cache gradients", detail, detail_size);
  if (status != LOTUS_HIP_PASS) return status;
  const float inv = 1.0f / static_cast<float>(count);
  if (weight_grad)
    for (int i = 0; i < kCacheParams; ++i) weight_grad[i] = wg[i] * inv;
  if (grid_grad)
    for (uint64_t i = 0; i < c.grid_entries; ++i) grid_grad[i] *= inv;
  if (loss) *loss = loss_sum / (3.0f * static_cast<float>(count));
  return LOTUS_HIP_PASS;
}

int lotus_hip_cache_train(lotus_hip_renderer* r, const void* samples_device, uint32_t count, float* loss,
                          char* detail, size_t detail_size) {
  if (!cache_usable(r, detail, detail_size)) return LOTUS_HIP_FAIL;
  if (count == 0 || count > r->cache.config.batch_capacity || !samples_device) {
    set_detail(detail, detail_size, "count must lie in [1, batch_capacity] with a device sample buffer");
    return LOTUS_HIP_FAIL;
  }
  CacheState& c = r->cache;
  launch_fill1(c.loss, 1, 0.0f, r->stream);
  cache_backprop(c, static_cast<const float*>(samples_device), count, nullptr, 0, 1, r->stream);
  cache_adam(c, count, nullptr, 0, 1, r->stream);
  float loss_sum = 0.0f;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&loss_sum, c.loss, sizeof(float), hipMemcpyDeviceToHost, r->stream);
  const int status = cache_finish(r, e, "cache training", detail, detail_size);
  if (status == LOTUS_HIP_PASS && loss) *loss = loss_sum / (3.0f * static_cast<float>(count));
  return status;
}

int lotus_hip_cache_parameters(lotus_hip_renderer* r, float* weights, float* grid, int write, char* detail,
                               size_t detail_size) {
  if (!cache_usable(r, detail, detail_size)) return LOTUS_HIP_FAIL;
  CacheState& c = r->cache;
  const hipMemcpyKind kind = write ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  hipError_t e = hipSuccess;
  if (weights)
    e = write ? hipMemcpyAsync(c.w32, weights, sizeof(float) * kCacheParams, kind, r->stream)
              : hipMemcpyAsync(weights, c.w32, sizeof(float) * kCacheParams, kind, r->stream);
  if (e == hipSuccess && grid)
    e = write ? hipMemcpyAsync(c.grid32, grid, sizeof(float) * c.grid_entries, kind, r->stream)
              : hipMemcpyAsync(grid, c.grid32, sizeof(float) * c.grid_entries, kind, r->stream);
  if (e == hipSuccess && write) {
    cache_reset_moments(c, r->stream);
    cache_sync_half(c, r->stream);
    r->key_valid = false;
  }
  return cache_finish(r, e, "cache parameters", detail, detail_size);
}

}  // extern "C"

namespace {


// The camera's world-space position: the view matrix's inverse applied to
// the origin, in double (Gauss-Jordan with partial pivoting).
void camera_eye(const float* view, float* eye) {
  double a[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      a[r][c] = view[c * 4 + r];
      a[r][4 + c] = r == c ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int pivot = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[pivot][c])) pivot = r;
    if (a[pivot][c] == 0.0) {
      eye[0] = eye[1] = eye[2] = 0.0f;
      return;
    }
    for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[pivot][k]);
    const double d = a[c][c];
    for (int k = 0; k < 8; ++k) a[c][k] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
    }
  }
  // column 3 of the inverse, divided by its w
  for (int r = 0; r < 3; ++r) eye[r] = static_cast<float>(a[r][4 + 3] / a[3][4 + 3]);
}

void default_validation(const float* in, float* out) {
  const bool given = in[0] != 0.0f || in[1] != 0.0f || in[2] != 0.0f || in[3] != 0.0f;
  const float defaults[4] = {0.01f, 0.01f, 0.9f, 32.0f};  // temporal.hpp:28-33
  std::memcpy(out, given ? in : defaults, 16);
}

}  // namespace

extern "C" {

int lotus_hip_render_temporal(lotus_hip_renderer* r, const lotus_hip_frame_params* P,
                              lotus_hip_temporal_settings* ts, lotus_hip_frame_result* res, char* detail,
                              size_t detail_size) {
  if (!r || !P || !ts) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  // TemporalFrameSettings rules (temporal.hpp:58-103, vulkan_backend.cpp:1066-1076)
  if (P->output != LOTUS_HIP_OUTPUT_RADIANCE) {
    set_detail(detail, detail_size, "temporal frames render radiance");
    return LOTUS_HIP_FAIL;
  }
  if (ts->accumulate_radiance && (ts->max_history_samples == 0 || ts->max_history_samples > 65536)) {
    set_detail(detail, detail_size, "max_history_samples must be in [1, 65536]");
    return LOTUS_HIP_FAIL;
  }
  if (ts->filter_iterations > 8 || (ts->filter_iterations != 0 && !ts->accumulate_radiance)) {
    set_detail(detail, detail_size, "filter_iterations needs accumulate_radiance and is at most 8");
    return LOTUS_HIP_FAIL;
  }
  if (ts->settle && !ts->accumulate_radiance) {  // vulkan_backend.cpp:1750-1751
    set_detail(detail, detail_size, "settling requires radiance accumulation");
    return LOTUS_HIP_FAIL;
  }
  // vulkan_backend.cpp:1752-1755
  if (ts->history_rectification != 0 &&
      (!ts->accumulate_radiance || ts->history_rectification > LOTUS_HIP_RECTIFY_REFLECTION)) {
    set_detail(detail, detail_size, "history rectification requires radiance accumulation and a known mode");
    return LOTUS_HIP_FAIL;
  }
  const bool restir = P->light_sampling == LOTUS_HIP_LIGHT_RESTIR;
  if (ts->restir_temporal && (!restir || P->frame_count != 1 || ts->restir_max_history_candidates == 0 ||
                              ts->restir_max_history_candidates > 65536)) {
    set_detail(detail, detail_size,
               "temporal reservoirs require RenderTemporalScene, RestirInitial, one sample and history candidates in [1, 65536]");
    return LOTUS_HIP_FAIL;
  }
  if (ts->restir_spatial && (!restir || P->frame_count != 1 || ts->restir_max_history_candidates == 0 ||
                             ts->restir_max_history_candidates > 65536 || ts->restir_spatial_radius == 0 ||
                             ts->restir_spatial_radius > 32 ||
                             ts->restir_spatial_neighbours > LOTUS_HIP_SPATIAL_DISK)) {
    set_detail(detail, detail_size,
               "spatial reservoirs require RenderTemporalScene, RestirInitial, one sample, history candidates in [1, 65536] and radius in [1, 32]");
    return LOTUS_HIP_FAIL;
  }
  auto& t = r->temporal;
  if (P->width == 0 || P->height == 0 || !std::isfinite(P->display_window[0]) ||
      !std::isfinite(P->display_window[1]) || !std::isfinite(P->display_window[2]) ||
      !std::isfinite(P->display_window[3])) {
    set_detail(detail, detail_size, "temporal frame extent must be nonempty");
    return LOTUS_HIP_FAIL;
  }
  if (t.previous && ts->frame_index <= t.frame_index) {
    set_detail(detail, detail_size, "temporal frame indices must increase");
    return LOTUS_HIP_FAIL;
  }
  // TemporalHistory::Advance (temporal.cpp:114-147): why history is dropped
  uint32_t reason = 0;
  if (!t.previous) reason = 1;                                    // FirstFrame
  else if (ts->camera_cut) reason = 3;                            // CameraCut
  else if (ts->frame_index - t.frame_index != 1) reason = 4;      // FrameGap
  else if (P->width != t.width || P->height != t.height ||         // TargetChanged
           std::memcmp(P->display_window, t.raw_display, 16) != 0 || std::memcmp(P->data_window, t.raw_data, 16) != 0)
    reason = 5;
  else if (ts->history_key != t.history_key) reason = 6;          // HistoryKeyChanged
  // vulkan_backend.cpp:1770-1789: a frame with a fresh batch drops the history
  // its transport settings no longer match (RenderSettingsChanged), or a
  // scene change temporal history cannot carry: an incompatible one, or any
  // while reservoirs reuse light samples of emissive triangles (SceneChanged)
  const bool reuse = restir && (ts->restir_temporal || ts->restir_spatial);
  const bool fresh = ts->accumulate_radiance || reuse;
  std::array<uint32_t, 14> transport{P->light_sampling, P->restir_proposal, P->restir_candidates,
                                     ts->restir_temporal ? 1u : 0u, ts->restir_max_history_candidates,
                                     ts->restir_spatial ? 1u : 0u, ts->restir_spatial_radius,
                                     ts->accumulate_radiance ? 1u : 0u, P->sample_index, P->max_bounces};
  std::memcpy(transport.data() + 10, P->clear_color, 16);
  if (reason == 0 && fresh) {
    if (!t.transport_valid || transport != t.transport)
      reason = 7;
    else if (r->scene_generation != t.scene_generation &&
             (r->scene_epoch != t.scene_epoch || (reuse && r->emissive_triangles != 0)))
      reason = 8;
  }
  (void)hipSetDevice(r->device);
  hipError_t e = ensure_temporal(r, P->width, P->height);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "temporal allocation", e));
    return LOTUS_HIP_FAIL;
  }
  const bool available = reason == 0 && t.previous;
  const uint32_t cur = t.cur, prev = 1u - t.cur;
  const uint32_t rcur = t.rcur, rprev = 1u - t.rcur;
  // vulkan_backend.cpp:1790-1795: settling needs nothing this frame depends
  // on to differ from the preceding frame, which left its colour as the
  // prior; reservoir reuse traces a fresh batch every frame and never settles
  float clip[16];
  vulkan_world_to_clip(P->world_to_clip, clip);
  const bool settling = ts->settle && !(restir && (ts->restir_temporal || ts->restir_spatial)) && t.prior_valid &&
                        available && r->scene_generation == t.scene_generation &&
                        std::memcmp(ts->view, t.view, 64) == 0 && std::memcmp(clip, t.world_to_clip, 64) == 0;
  // vulkan_backend.cpp:1805-1813: a fresh batch continues the stream's
  // samples; a settling frame continues the still accumulation from the
  // host's sample index
  const uint32_t sample_offset = available ? t.sample_offset : 0u;
  const uint32_t settled_before = settling ? r->accumulated : 0u;
  lotus_hip_frame_params B = *P;
  if (fresh && !settling) B.sample_index += sample_offset;
  TemporalFrameState tf;
  std::memcpy(tf.view, ts->view, 64);
  TemporalConst& T = tf.T;
  std::memcpy(T.current_view, ts->view, 64);
  std::memcpy(T.previous_view, t.view, 64);
  std::memcpy(T.previous_world_to_clip, t.world_to_clip, 64);
  std::memcpy(T.previous_display, t.display, 16);
  std::memcpy(T.previous_data, t.data, 16);
  default_validation(ts->validation, T.validation);
  T.available = available ? 1u : 0u;
  T.width = P->width;
  T.height = P->height;
  T.fresh_samples = std::max(P->frame_count, 1u);
  T.cap = ts->accumulate_radiance ? ts->max_history_samples : 1u;
  T.iterations = ts->accumulate_radiance ? ts->filter_iterations : 0u;
  T.rectification = ts->history_rectification;
  camera_eye(ts->view, T.current_eye);
  std::memcpy(T.previous_eye, t.eye, sizeof t.eye);
  // the per-instance identities are the caller's host arrays: copied to the
  // device, which the surface pass and the reuse passes read them on
  tf.I = TemporalInstances{nullptr, nullptr, nullptr, nullptr};
  if (ts->previous_from_current || ts->instance_ids || ts->material_ids || ts->instance_flags) {
    const size_t n = r->view.instance_count;
    const size_t xf_bytes = 64 * n, id_bytes = 8 * n, flag_bytes = 4 * n;
    const size_t need = xf_bytes + 2 * id_bytes + flag_bytes;
    if (need > r->temporal_ids_bytes) {
      dev_free(r->temporal_ids);
      r->temporal_ids_bytes = 0;
      e = dev_alloc(r->temporal_ids, need);
      if (e == hipSuccess) r->temporal_ids_bytes = need;
    }
    unsigned char* d = r->temporal_ids;
    auto upload = [&](const void* host, size_t offset, size_t bytes) -> void* {
      if (!host || e != hipSuccess) return nullptr;
      e = hipMemcpyAsync(d + offset, host, bytes, hipMemcpyHostToDevice, r->stream);
      return d + offset;
    };
    if (e == hipSuccess && n != 0) {
      tf.I.previous_from_current = static_cast<const float*>(upload(ts->previous_from_current, 0, xf_bytes));
      tf.I.instance_id = static_cast<const uint64_t*>(upload(ts->instance_ids, xf_bytes, id_bytes));
      tf.I.material_id = static_cast<const uint64_t*>(upload(ts->material_ids, xf_bytes + id_bytes, id_bytes));
      tf.I.flags = static_cast<const uint32_t*>(upload(ts->instance_flags, xf_bytes + 2 * id_bytes, flag_bytes));
    }
    if (e != hipSuccess) {
      set_detail(detail, detail_size, latch(r, "temporal identity upload", e));
      return LOTUS_HIP_FAIL;
    }
  }
  tf.surface = t.surf[cur];
  tf.previous_surface = t.surf[prev];
  tf.validity = t.validity;
  tf.radiance = t.rad[rcur];
  tf.old_radiance = t.rad[rprev];
  tf.filter = t.filter;
  tf.accumulate = ts->accumulate_radiance != 0;
  tf.fresh = fresh;
  tf.settling = settling;
  tf.prior = t.prior;
  tf.max_samples = P->max_samples;
  tf.batch_host = ts->batch;
  tf.reuse = restir && (ts->restir_temporal || ts->restir_spatial);
  tf.spatial = restir && ts->restir_spatial;
  RestirReuseArgs& A = tf.A;
  A.surf = t.rsurf[cur];
  A.prev_surf = t.rsurf[prev];
  A.prev_hist = t.rhist[prev];
  A.temporal_out = tf.spatial ? t.rspin : t.rhist[cur];
  A.spatial_out = t.rhist[cur];
  A.spatial_sources = t.rsources;
  A.tsurf = t.surf[cur];
  A.validity = t.validity;
  A.history = ts->restir_temporal && available && t.reservoirs ? 1u : 0u;
  A.max_history = ts->restir_max_history_candidates;
  A.radius = ts->restir_spatial_radius;
  A.disk = ts->restir_spatial_neighbours == LOTUS_HIP_SPATIAL_DISK ? 1u : 0u;
  if (tf.accumulate && !(available && t.radiance))
    launch_fill4(reinterpret_cast<float4*>(t.rad[rprev]), 2u * P->width * P->height,
                 make_float4(0.0f, 0.0f, 0.0f, 0.0f), r->stream);
  const int status = render_impl(r, &B, res, detail, detail_size, &tf);
  if (status != LOTUS_HIP_PASS) return status;
  const size_t npix = size_t(P->width) * P->height;
  const uint32_t samples = r->accumulated;  // this frame's samples per pixel
  // the changing frame's colour becomes the next settling frames' prior
  if (ts->settle && !settling)
    e = hipMemcpy(t.prior, r->q.color, npix * sizeof(float4), hipMemcpyDeviceToDevice);
  if (e == hipSuccess && ts->surface)
    e = hipMemcpy(ts->surface, t.surf[cur], npix * sizeof(TemporalPixel), hipMemcpyDeviceToHost);
  if (e == hipSuccess && ts->validity)
    e = hipMemcpy(ts->validity, t.validity, npix * sizeof(TemporalValidityPixel), hipMemcpyDeviceToHost);
  // a settling frame writes no history: its product is the retained one
  if (e == hipSuccess && ts->radiance && tf.accumulate)
    e = hipMemcpy(ts->radiance, t.rad[settling ? rprev : rcur], npix * sizeof(TemporalRadiancePixel),
                  hipMemcpyDeviceToHost);
  // the reuse passes' records: the temporal pass writes the final records
  // alone, or the spatial pass's input when that pass follows
  if (e == hipSuccess && ts->restir_temporal_records && restir && ts->restir_temporal)
    e = hipMemcpy(ts->restir_temporal_records, tf.spatial ? t.rspin : t.rhist[cur], npix * 4 * sizeof(float4),
                  hipMemcpyDeviceToHost);
  if (e == hipSuccess && ts->restir_spatial_records && tf.spatial)
    e = hipMemcpy(ts->restir_spatial_records, t.rhist[cur], npix * 4 * sizeof(float4), hipMemcpyDeviceToHost);
  if (e == hipSuccess && ts->restir_spatial_sources && tf.spatial)
    e = A.disk ? hipMemcpy(ts->restir_spatial_sources, t.rsources, npix * 2 * sizeof(uint4), hipMemcpyDeviceToHost)
               : hipMemcpy2D(ts->restir_spatial_sources, sizeof(uint4), t.rsources, 2 * sizeof(uint4),
                             sizeof(uint4), npix, hipMemcpyDeviceToHost);
  if (e == hipSuccess && ts->restir_surfaces && tf.reuse)
    e = hipMemcpy(ts->restir_surfaces, t.rsurf[cur], npix * kReuseSurfaceWords * sizeof(float4), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "temporal readback", e));
    return LOTUS_HIP_FAIL;
  }
  t.previous = true;
  t.frame_index = ts->frame_index;
  t.history_key = ts->history_key;
  t.scene_generation = r->scene_generation;
  t.scene_epoch = r->scene_epoch;
  std::memcpy(t.view, ts->view, 64);
  std::memcpy(t.eye, T.current_eye, sizeof t.eye);
  vulkan_world_to_clip(P->world_to_clip, t.world_to_clip);
  resolved_windows(P, t.display, t.data);
  std::memcpy(t.raw_display, P->display_window, 16);
  std::memcpy(t.raw_data, P->data_window, 16);
  t.transport = transport;
  t.transport_valid = fresh;
  t.light_sampling = P->light_sampling;
  t.max_bounces = P->max_bounces;
  t.candidates = P->restir_candidates;
  t.proposal = P->restir_proposal;
  t.reservoirs = tf.reuse;
  t.radiance = tf.accumulate;
  t.cur = prev;
  // vulkan_backend.cpp:2085-2097
  if (!settling) {
    t.sample_offset = fresh ? sample_offset + samples : 0u;
    t.rcur = rprev;
  }
  if (fresh && !settling) {  // the next progressive mean starts over
    r->key_valid = false;
    r->accumulated = 0;
  }
  t.prior_valid = ts->settle != 0;
  ts->reset_reason = reason;
  ts->history_available = available ? 1u : 0u;
  ts->settling = settling ? 1u : 0u;
  ts->settled_samples = settling ? samples : 0u;
  ts->batch_samples = samples - std::min(samples, settled_before);
  ts->batch_sample_index = B.sample_index + settled_before;
  return LOTUS_HIP_PASS;
}

}  // extern "C"

namespace {

// EstimateTemporalVariance and FilterTemporalRadiance over caller device
// buffers with `filter` (3 float4 per pixel) as the scratch.
int svgf_run(lotus_hip_renderer* r, const void* surface, void* radiance, uint32_t width, uint32_t height,
             uint32_t batch_samples, uint32_t max_history_samples, uint32_t iterations, void* color, float4* filter,
             char* detail, size_t detail_size) {
  TemporalConst T{};
  T.width = width;
  T.height = height;
  T.fresh_samples = batch_samples;
  T.cap = max_history_samples;
  T.iterations = iterations;
  const Launch L = launch_for(r, r->stream);
  auto* surf = static_cast<const TemporalPixel*>(surface);
  auto* rad = static_cast<TemporalRadiancePixel*>(radiance);
  launch_temporal_variance(L, T, surf, rad, filter);
  if (iterations != 0) launch_temporal_filter(L, T, surf, rad, filter, static_cast<float4*>(color));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "filter", e));
    return LOTUS_HIP_FAIL;
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

bool svgf_usable(lotus_hip_renderer* r, const void* surface, const void* radiance, uint32_t width, uint32_t height,
                 uint32_t batch_samples, uint32_t max_history_samples, uint32_t iterations, const void* color,
                 char* detail, size_t detail_size) {
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return false;
  }
  if (!surface || !radiance || width == 0 || height == 0 || batch_samples == 0 || max_history_samples == 0 ||
      iterations > 8 || (iterations != 0 && !color) || uint64_t(width) * height > 0xFFFFFFFFull / 3) {
    set_detail(detail, detail_size, "invalid filter input");
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

int lotus_hip_svgf(lotus_hip_renderer* r, const void* surface, void* radiance, uint32_t width, uint32_t height,
                   uint32_t batch_samples, uint32_t max_history_samples, uint32_t iterations, void* color,
                   char* detail, size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!svgf_usable(r, surface, radiance, width, height, batch_samples, max_history_samples, iterations, color, detail,
                   detail_size))
    return LOTUS_HIP_FAIL;
  (void)hipSetDevice(r->device);
  const uint32_t P = width * height;
  if (r->svgf_pixels < P) {
    dev_free(r->svgf_filter);
    r->svgf_pixels = 0;
    const hipError_t e = dev_alloc(r->svgf_filter, 4ull * P);
    if (e != hipSuccess) {
      set_detail(detail, detail_size, latch(r, "filter", e));
      return LOTUS_HIP_FAIL;
    }
    r->svgf_pixels = P;
  }
  return svgf_run(r, surface, radiance, width, height, batch_samples, max_history_samples, iterations, color,
                  r->svgf_filter, detail, detail_size);
}

int lotus_hip_accumulate(lotus_hip_renderer* r, const void* batch, const void* surface, const void* validity,
                         const void* old_radiance, uint32_t width, uint32_t height, uint32_t fresh_samples,
                         uint32_t max_history_samples, uint32_t rectification, void* radiance, char* detail,
                         size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  if (!batch || !surface || !validity || !old_radiance || !radiance || width == 0 || height == 0 ||
      fresh_samples == 0 || max_history_samples == 0 || max_history_samples > 65536 ||
      rectification > LOTUS_HIP_RECTIFY_REFLECTION || uint64_t(width) * height > 0xFFFFFFFFull) {
    set_detail(detail, detail_size, "invalid accumulation input");
    return LOTUS_HIP_FAIL;
  }
  (void)hipSetDevice(r->device);
  TemporalConst T{};
  T.width = width;
  T.height = height;
  T.fresh_samples = fresh_samples;
  T.cap = max_history_samples;
  T.rectification = rectification;
  const Launch L = launch_for(r, r->stream);
  launch_temporal_accumulate(L, T, static_cast<const float4*>(batch), static_cast<const TemporalPixel*>(surface),
                             static_cast<const TemporalValidityPixel*>(validity),
                             static_cast<const TemporalRadiancePixel*>(old_radiance),
                             static_cast<TemporalRadiancePixel*>(radiance));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "accumulation", e));
    return LOTUS_HIP_FAIL;
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

int lotus_hip_svgf_static(lotus_hip_renderer* r, const void* surface, void* radiance, uint32_t width,
                          uint32_t height, uint32_t batch_samples, uint32_t max_history_samples, uint32_t iterations,
                          void* color, char* detail, size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!svgf_usable(r, surface, radiance, width, height, batch_samples, max_history_samples, iterations, color, detail,
                   detail_size))
    return LOTUS_HIP_FAIL;
  if (uint64_t(width) * height > kStaticSvgfPixels) {
    set_detail(detail, detail_size,
               "the static filter scratch holds at most " + std::to_string(kStaticSvgfPixels) + " pixels, not " +
                   std::to_string(uint64_t(width) * height));
    return LOTUS_HIP_FAIL;
  }
  if (!r->static_scratch) {
    set_detail(detail, detail_size, r->static_scratch_error);
    return LOTUS_HIP_FAIL;
  }
  (void)hipSetDevice(r->device);
  return svgf_run(r, surface, radiance, width, height, batch_samples, max_history_samples, iterations, color,
                  r->static_scratch, detail, detail_size);
}

int lotus_hip_write_render_buffer(lotus_hip_renderer* r, const void* product, uint32_t product_format,
                                  uint32_t source_width, uint32_t source_height, void* buffer, uint32_t buffer_format,
                                  uint32_t buffer_width, uint32_t buffer_height, char* detail, size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  const bool colour_pair = (product_format == LOTUS_HIP_BUFFER_FLOAT32X4 && buffer_format == LOTUS_HIP_BUFFER_UNORM8X4) ||
                           (product_format == LOTUS_HIP_BUFFER_UNORM8X4 && buffer_format == LOTUS_HIP_BUFFER_FLOAT32X4);
  if (buffer_pixel_bytes(product_format) == 0 || buffer_pixel_bytes(buffer_format) == 0 ||
      (product_format != buffer_format && !colour_pair)) {
    set_detail(detail, detail_size, "the product's format does not convert to the render buffer's");
    return LOTUS_HIP_FAIL;
  }
  if (source_width != buffer_width || source_height != buffer_height) {
    set_detail(detail, detail_size, "the product's size differs from the render buffer's");
    return LOTUS_HIP_FAIL;
  }
  if (source_width == 0 || source_height == 0) {
    set_detail(detail, detail_size, "");
    return LOTUS_HIP_PASS;
  }
  if (!product || !buffer) {
    set_detail(detail, detail_size, "invalid render buffer input");
    return LOTUS_HIP_FAIL;
  }
  (void)hipSetDevice(r->device);
  launch_write_rows_flipped(product, product_format, buffer, buffer_format, source_width, source_height, r->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "render buffer write", e));
    return LOTUS_HIP_FAIL;
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

int lotus_hip_write_ids(lotus_hip_renderer* r, void* buffer, uint32_t buffer_format, uint32_t width, uint32_t height,
                        int32_t value, char* detail, size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  if (buffer_format != LOTUS_HIP_BUFFER_INT32) {
    set_detail(detail, detail_size, "ids are written to Int32 render buffers only");
    return LOTUS_HIP_FAIL;
  }
  const uint64_t count = uint64_t(width) * height;
  if (count != 0 && !buffer) {
    set_detail(detail, detail_size, "invalid render buffer input");
    return LOTUS_HIP_FAIL;
  }
  (void)hipSetDevice(r->device);
  launch_fill_ids(static_cast<int32_t*>(buffer), count, value, r->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "id write", e));
    return LOTUS_HIP_FAIL;
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

namespace {
// The checks WriteRowsFlippedLocked makes before writing (adapter.cpp:
// 2713-2720): the buffer's format is the AOV's and its size the source's.
bool aov_target_ok(uint32_t aov_format, uint32_t source_width, uint32_t source_height, const void* buffer,
                   uint32_t buffer_format, uint32_t buffer_width, uint32_t buffer_height, char* detail,
                   size_t detail_size) {
  if (buffer_format != aov_format) {
    set_detail(detail, detail_size, "the AOV's format differs from the render buffer's");
    return false;
  }
  if (source_width != buffer_width || source_height != buffer_height) {
    set_detail(detail, detail_size, "the product's size differs from the render buffer's");
    return false;
  }
  if (uint64_t(source_width) * source_height != 0 && !buffer) {
    set_detail(detail, detail_size, "invalid render buffer input");
    return false;
  }
  return true;
}

int finish_aov_write(lotus_hip_renderer* r, const char* what, char* detail, size_t detail_size) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, what, e));
    return LOTUS_HIP_FAIL;
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}
}  // namespace

int lotus_hip_write_temporal_aov(lotus_hip_renderer* r, uint32_t aov, const void* surface, const void* validity,
                                 const void* radiance, uint32_t source_width, uint32_t source_height, void* buffer,
                                 uint32_t buffer_format, uint32_t buffer_width, uint32_t buffer_height, char* detail,
                                 size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  if (aov > LOTUS_HIP_AOV_HISTORY_REJECTION) {
    set_detail(detail, detail_size, "not a temporal AOV");
    return LOTUS_HIP_FAIL;
  }
  if (!aov_target_ok(temporal_aov_format(aov), source_width, source_height, buffer, buffer_format, buffer_width,
                     buffer_height, detail, detail_size))
    return LOTUS_HIP_FAIL;
  if (uint64_t(source_width) * source_height == 0) {
    set_detail(detail, detail_size, "");
    return LOTUS_HIP_PASS;
  }
  if ((aov == LOTUS_HIP_AOV_MOTION && !surface) ||
      ((aov == LOTUS_HIP_AOV_HISTORY_VALIDITY || aov == LOTUS_HIP_AOV_HISTORY_REJECTION) && !validity)) {
    set_detail(detail, detail_size, "the AOV's temporal records are missing");
    return LOTUS_HIP_FAIL;
  }
  (void)hipSetDevice(r->device);
  launch_temporal_aov(aov, surface, validity, radiance, buffer, source_width, source_height, r->stream);
  return finish_aov_write(r, "temporal AOV write", detail, detail_size);
}

int lotus_hip_write_restir_aov(lotus_hip_renderer* r, uint32_t aov, const void* initial, const void* resolved,
                               uint32_t source_width, uint32_t source_height, void* buffer, uint32_t buffer_format,
                               uint32_t buffer_width, uint32_t buffer_height, char* detail, size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  if (aov > LOTUS_HIP_AOV_RESTIR_DIRECT) {
    set_detail(detail, detail_size, "not a ReSTIR AOV");
    return LOTUS_HIP_FAIL;
  }
  if (!aov_target_ok(restir_aov_format(aov), source_width, source_height, buffer, buffer_format, buffer_width,
                     buffer_height, detail, detail_size))
    return LOTUS_HIP_FAIL;
  if (uint64_t(source_width) * source_height == 0) {
    set_detail(detail, detail_size, "");
    return LOTUS_HIP_PASS;
  }
  (void)hipSetDevice(r->device);
  launch_restir_aov(aov, initial, resolved, buffer, source_width, source_height, r->stream);
  return finish_aov_write(r, "ReSTIR AOV write", detail, detail_size);
}

int lotus_hip_reproject(lotus_hip_renderer* r, const void* current, const void* previous, uint32_t width,
                        uint32_t height, const float* validation, void* validity, char* detail, size_t detail_size) {
  if (!r) return LOTUS_HIP_FAIL;
  if (!r->failure.empty()) {
    set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
    return LOTUS_HIP_FAIL;
  }
  if (!current || !validity || width == 0 || height == 0) {
    set_detail(detail, detail_size, "invalid reprojection input");
    return LOTUS_HIP_FAIL;
  }
  (void)hipSetDevice(r->device);
  TemporalConst T{};
  const float zero[4] = {0, 0, 0, 0};
  default_validation(validation ? validation : zero, T.validation);
  T.available = previous ? 1u : 0u;
  T.width = width;
  T.height = height;
  const Launch L = launch_for(r, r->stream);
  launch_temporal_reproject(L, T, static_cast<const TemporalPixel*>(current), static_cast<const TemporalPixel*>(previous),
                            static_cast<TemporalValidityPixel*>(validity));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(r, "reprojection", e));
    return LOTUS_HIP_FAIL;
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

}  // extern "C"

namespace {

// The bookkeeping of a deferred sample once its batch has completed: ray
// counts, the next sample's rounds, the accumulation count, the colour.
// What a graphed batch's captured launches of one view depend on besides
// its frame variables: the parameters (but the sample index), the scene's
// generation, the launch tuning and the view's buffers.
std::vector<uint8_t> graph_signature(const lotus_hip_renderer* r, const lotus_hip_frame_params* P) {
  lotus_hip_frame_params p = *P;
  p.sample_index = 0;
  std::vector<uint8_t> out(sizeof p + sizeof r->scene_generation + sizeof r->tuning + 4 * sizeof(void*));
  uint8_t* o = out.data();
  std::memcpy(o, &p, sizeof p);
  o += sizeof p;
  std::memcpy(o, &r->scene_generation, sizeof r->scene_generation);
  o += sizeof r->scene_generation;
  std::memcpy(o, &r->tuning, sizeof r->tuning);
  o += sizeof r->tuning;
  const void* buffers[4] = {r->q.thr, r->q.color, r->counters, r->stats};
  std::memcpy(o, buffers, sizeof buffers);
  return out;
}

// A replayed view's host side of render_impl's deferred sample: the
// accumulation's restart rule, and the frame variables of a sample that
// adds (false when none would: the captured launches no longer apply).
bool replay_prelude(lotus_hip_renderer* r, const lotus_hip_frame_params* P, uint32_t* vars) {
  float vk[16];
  vulkan_world_to_clip(P->world_to_clip, vk);
  restart_if_changed(r, P, vk, false, false);
  if (P->max_samples != 0 && r->accumulated >= P->max_samples) return false;
  if (!r->graph_deferred.add) return false;
  vars[0] = P->sample_index + r->accumulated;
  vars[1] = r->accumulated;
  vars[2] = 1u;
  vars[3] = 0u;
  r->deferred = r->graph_deferred;
  r->deferred.pending = true;
  return true;
}

// The readbacks of a completed batch, every view's counters and colour
// enqueued on the leader's stream with one synchronisation for all.
hipError_t batch_readback(lotus_hip_renderer* const* renderers, const lotus_hip_frame_params* params,
                          lotus_hip_frame_result* results, uint32_t count, hipStream_t stream,
                          std::vector<std::vector<uint32_t>>& hosts) {
  hosts.resize(count);
  hipError_t e = hipSuccess;
  for (uint32_t i = 0; i < count && e == hipSuccess; ++i) {
    lotus_hip_renderer* r = renderers[i];
    hosts[i].assign(frame_counter_words(r->deferred.stride), 0u);
    e = hipMemcpyAsync(hosts[i].data(), r->counters, hosts[i].size() * 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess && results && results[i].color)
      e = hipMemcpyAsync(results[i].color, r->q.color, 16ull * params[i].width * params[i].height,
                         hipMemcpyDeviceToHost, stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  return e;
}

int batch_finish(lotus_hip_renderer* r, const lotus_hip_frame_params* P, lotus_hip_frame_result* res, double gpu_ms,
                 const std::vector<uint32_t>& host) {
  auto& d = r->deferred;
  d.pending = false;
  std::vector<uint32_t> counts;
  uint64_t ext = 0, sh = 0, round_ext = 0, round_sh = 0;
  if (d.add) {
    counts.assign(d.stride, 0u);
    size_t last = 0;
    for (uint32_t b = 0; b < d.stride; ++b) {
      counts[b] = host[counter_word(b, 0)];
      if (counts[b]) last = b;
      if (b >= 1) ext += host[counter_word(b, 0)];
      sh += host[counter_word(b, 2)];
      if (b >= 1 && b < d.rounds) round_ext += host[counter_word(b, 0)];
      if (b + 1 < d.rounds || (b < d.rounds && d.rounds + 1 >= d.stride)) round_sh += host[counter_word(b, 2)];
    }
    counts.resize(host[0] ? last + 1 : 0);
    r->rounds_hint = wavefront_rounds(counts, r->tuning.tail_divisor);
    ++r->accumulated;
  }
  if (res) {
    res->path_counts_size = static_cast<uint32_t>(std::min<size_t>(counts.size(), res->path_counts_capacity));
    if (res->path_counts) std::memcpy(res->path_counts, counts.data(), res->path_counts_size * 4ull);
    res->wavefront_rounds = counts.empty() ? 0 : d.rounds;
    res->samples_per_pixel = r->accumulated;
    res->frames_rendered = 1;
    res->camera_rays = d.add ? host[0] : 0;
    res->extension_rays = ext;
    res->shadow_rays = sh + host[counter_word(0, 0) + kCounterSpread / 2];  // + parallel_ray_count
    res->trace_launches = d.launches;
    res->traversal_rays[0] = d.add ? host[0] : 0;
    res->traversal_rays[1] = round_ext;
    res->traversal_rays[2] = round_sh;
    res->traversal_rays[3] = (ext - round_ext) + (sh - round_sh);
    res->frame_gpu_ms = gpu_ms;
  }
  return LOTUS_HIP_PASS;
}

}  // namespace

extern "C" {

// A fused batch's paths, enqueued on the leader's stream once every view's
// stream has joined it: the views whose frames the megakernel runs (those
// render_impl left pending) go into one device table, grouped by kernel
// variant (plain, extras, covered surfaces), each group one launch fed by
// its own work-queue counter; then each such view's resolve on its own
// stream, joined back. Views render_impl ran in full are already done.
hipError_t fused_paths(lotus_hip_renderer* root, lotus_hip_renderer* const* renderers, uint32_t count) {
  std::vector<PathView> groups[3];
  size_t need = 4;  // spill entries per lane: the deepest view's (as ensure_launch sizes them)
  for (uint32_t i = 0; i < count; ++i) {
    lotus_hip_renderer* r = renderers[i];
    if (!r->fused_pending) continue;
    const SceneView& S = r->fused_view.S;
    const int g = S.covered_instances != 0 ? 2 : scene_has_extras(S) ? 1 : 0;
    PathView v = r->fused_view;
    const uint32_t first = groups[g].empty() ? 0u : groups[g].back().first_chunk + path_chunks(groups[g].back().F);
    v.first_chunk = first;
    groups[g].push_back(v);
    const size_t stack = 3 * size_t(r->scene.depth) + 4;
    need = std::max(need, stack > size_t(kRing) ? stack - kRing : size_t(0));
  }
  std::vector<PathView> table;
  for (const auto& g : groups) table.insert(table.end(), g.begin(), g.end());
  hipError_t e = hipSuccess;
  if (!table.empty()) {
    if (root->fused_table_count < table.size()) {
      dev_free(root->fused_table);
      root->fused_table_count = 0;
      e = dev_alloc(root->fused_table, table.size());
      if (e == hipSuccess) root->fused_table_count = static_cast<uint32_t>(table.size());
    }
    if (e == hipSuccess && !root->fused_queue) e = dev_alloc(root->fused_queue, 3 * kXcds);
    if (e == hipSuccess && (!root->fused_spill || root->fused_spill_entries < need)) {
      dev_free(root->fused_spill);
      root->fused_spill_entries = 0;
      e = dev_alloc(root->fused_spill, need * size_t(root->grid.max_trace_lanes));
      if (e == hipSuccess) root->fused_spill_entries = need;
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(root->fused_table, table.data(), table.size() * sizeof(PathView), hipMemcpyHostToDevice,
                         root->stream);
    if (e == hipSuccess) launch_zero(root->fused_queue, 3 * kXcds, root->stream);
    Launch L = launch_for(root, root->stream);
    L.grid = root->grid;  // the batch's launch: not the leader view's resident share
    L.spill = root->fused_spill;
    L.spill_cap = static_cast<int>(root->fused_spill_entries);
    size_t offset = 0;
    for (int g = 0; g < 3 && e == hipSuccess; ++g) {
      if (groups[g].empty()) continue;
      const uint32_t chunks = groups[g].back().first_chunk + path_chunks(groups[g].back().F);
      if (chunks != 0)
        launch_paths_views(L, root->fused_table + offset, static_cast<uint32_t>(groups[g].size()), g >= 1, g == 2,
                           root->fused_queue + kXcds * g, chunks);
      offset += groups[g].size();
    }
    if (e == hipSuccess) e = hipGetLastError();
  }
  // every view's resolve after the paths, each on its own stream
  if (e == hipSuccess) e = hipEventRecord(root->join, root->stream);
  for (uint32_t i = 0; i < count && e == hipSuccess; ++i) {
    lotus_hip_renderer* r = renderers[i];
    if (!r->fused_pending) continue;
    if (i != 0) e = hipStreamWaitEvent(r->stream, root->join, 0);
    if (e == hipSuccess) launch_resolve(launch_for(r, r->stream), r->fused_view.F, r->q, r->fused_add);
    r->fused_pending = false;
  }
  for (uint32_t i = 1; i < count && e == hipSuccess; ++i) {
    e = hipEventRecord(renderers[i]->join, renderers[i]->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(root->stream, renderers[i]->join, 0);
  }
  return e;
}

int lotus_hip_render_batch(lotus_hip_renderer* const* renderers, const lotus_hip_frame_params* params,
                           lotus_hip_frame_result* results, uint32_t count, uint32_t flags, double* gpu_ms,
                           char* detail, size_t detail_size) {
  if (!renderers || !params || count == 0) {
    set_detail(detail, detail_size, "a batch needs at least one view");
    return LOTUS_HIP_FAIL;
  }
  for (uint32_t i = 0; i < count; ++i) {
    lotus_hip_renderer* r = renderers[i];
    const lotus_hip_frame_params& P = params[i];
    if (!r) return LOTUS_HIP_FAIL;
    for (uint32_t j = 0; j < i; ++j)
      if (renderers[j] == r) {
        set_detail(detail, detail_size, "each view of a batch needs its own renderer");
        return LOTUS_HIP_FAIL;
      }
    if (!r->failure.empty()) {
      set_detail(detail, detail_size, "the renderer failed earlier: " + r->failure);
      return LOTUS_HIP_FAIL;
    }
    if (r->device != renderers[0]->device) {
      set_detail(detail, detail_size, "a batch runs on one device");
      return LOTUS_HIP_FAIL;
    }
    if (P.output != LOTUS_HIP_OUTPUT_RADIANCE || P.frame_count != 1 || P.light_sampling > LOTUS_HIP_LIGHT_MIS ||
        (P.flags & ~LOTUS_HIP_FRAME_SPLIT_KINDS) != 0) {
      set_detail(detail, detail_size, "batched views render one radiance sample with BSDF or MIS light sampling");
      return LOTUS_HIP_FAIL;
    }
  }
  const bool fused = (flags & LOTUS_HIP_BATCH_FUSED) != 0;
  if (fused && (flags & LOTUS_HIP_BATCH_GRAPH) != 0) {
    set_detail(detail, detail_size, "fused batches run eagerly, not as a graph");
    return LOTUS_HIP_FAIL;
  }
  lotus_hip_renderer* root = renderers[0];
  (void)hipSetDevice(root->device);
  hipError_t e = hipSuccess;
  // targets first: nothing allocates once the frame is being captured
  for (uint32_t i = 0; i < count && e == hipSuccess; ++i) {
    e = ensure_target(renderers[i], params[i].width, params[i].height);
    if (e == hipSuccess && !renderers[i]->join) e = hipEventCreateWithFlags(&renderers[i]->join, hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    set_detail(detail, detail_size, latch(root, "batch allocation", e));
    return LOTUS_HIP_FAIL;
  }
  const bool graph = (flags & LOTUS_HIP_BATCH_GRAPH) != 0;
  hipEvent_t start = root->events[root->events.size() - 2], stop = root->events[root->events.size() - 1];
  hipEvent_t fork = root->join;
  // a graphed batch's frame variables (4 words per view) in device memory
  if (graph && root->batch_vars_count < count) {
    dev_free(root->batch_vars);
    root->batch_vars_count = 0;
    e = dev_alloc(root->batch_vars, 4ull * count);
    if (e != hipSuccess) {
      set_detail(detail, detail_size, latch(root, "batch allocation", e));
      return LOTUS_HIP_FAIL;
    }
    root->batch_vars_count = count;
  }
  const std::vector<const lotus_hip_renderer*> views(renderers, renderers + count);
  std::vector<uint32_t> vars(4ull * count, 0u);
  std::vector<std::vector<uint8_t>> signatures(count);
  if (graph) {
    for (uint32_t i = 0; i < count; ++i) {
      renderers[i]->graph_vars = root->batch_vars + 4ull * i;
      signatures[i] = graph_signature(renderers[i], &params[i]);
    }
  }
  // Replay: the same views with the same parameters (but the sample index)
  // on the same scenes and buffers as the captured graph, each adding a
  // sample: the graph runs again with this frame's variables, uncaptured.
  bool replay = graph && root->batch_exec && root->batch_views == views;
  for (uint32_t i = 0; i < count && replay; ++i)
    replay = renderers[i]->graph_signature == signatures[i] && replay_prelude(renderers[i], &params[i], &vars[4ull * i]);
  if (replay) {
    e = hipMemcpyAsync(root->batch_vars, vars.data(), vars.size() * 4, hipMemcpyHostToDevice, root->stream);
    if (e == hipSuccess) e = hipEventRecord(start, root->stream);
    if (e == hipSuccess) e = hipGraphLaunch(root->batch_exec, root->stream);
    if (e == hipSuccess) e = hipEventRecord(stop, root->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(root->stream);
    if (e != hipSuccess) {
      for (uint32_t i = 0; i < count; ++i) renderers[i]->deferred.pending = false;
      set_detail(detail, detail_size, latch(root, "batched frame", e));
      return LOTUS_HIP_FAIL;
    }
    ++root->batch_replays;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, start, stop);
    if (gpu_ms) *gpu_ms = ms;
    std::vector<std::vector<uint32_t>> hosts;
    e = batch_readback(renderers, params, results, count, root->stream, hosts);
    if (e != hipSuccess) {
      for (uint32_t i = 0; i < count; ++i) renderers[i]->deferred.pending = false;
      set_detail(detail, detail_size, latch(root, "batch readback", e));
      return LOTUS_HIP_FAIL;
    }
    for (uint32_t i = 0; i < count; ++i) {
      (void)batch_finish(renderers[i], &params[i], results ? &results[i] : nullptr, ms, hosts[i]);
      if (results) results[i].batch_graph = 3u;
    }
    set_detail(detail, detail_size, "");
    return LOTUS_HIP_PASS;
  }
  if (graph) e = hipStreamBeginCapture(root->stream, hipStreamCaptureModeRelaxed);
  if (!graph && e == hipSuccess) e = hipEventRecord(start, root->stream);
  if (e == hipSuccess) e = hipEventRecord(fork, root->stream);
  for (uint32_t i = 1; i < count && e == hipSuccess; ++i) e = hipStreamWaitEvent(renderers[i]->stream, fork, 0);
  int status = LOTUS_HIP_PASS;
  for (uint32_t i = 0; i < count && e == hipSuccess && status == LOTUS_HIP_PASS; ++i) {
    renderers[i]->graph_batch = graph;
    renderers[i]->fuse = fused;
    renderers[i]->fused_pending = false;
    status = render_impl(renderers[i], &params[i], nullptr, detail, detail_size, nullptr, true);
    renderers[i]->graph_batch = false;
    renderers[i]->fuse = false;
    if (graph && status == LOTUS_HIP_PASS) {
      std::memcpy(&vars[4ull * i], renderers[i]->graph_host_vars, 16);
      renderers[i]->graph_deferred = renderers[i]->deferred;
    }
  }
  for (uint32_t i = 1; i < count && e == hipSuccess; ++i) {
    e = hipEventRecord(renderers[i]->join, renderers[i]->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(root->stream, renderers[i]->join, 0);
  }
  if (fused && e == hipSuccess && status == LOTUS_HIP_PASS) e = fused_paths(root, renderers, count);
  for (uint32_t i = 0; i < count; ++i) renderers[i]->fused_pending = false;
  hipGraph_t g = nullptr;
  bool updated_graph = false;
  if (graph) {
    const hipError_t end = hipStreamEndCapture(root->stream, &g);
    if (e == hipSuccess) e = end;
    if (e == hipSuccess && status == LOTUS_HIP_PASS) {
      // the same views as the last batch: the executable graph takes this
      // frame's parameters in place when the topology is unchanged
      const std::vector<const lotus_hip_renderer*> views(renderers, renderers + count);
      bool updated = false;
      if (root->batch_exec && root->batch_views == views) {
        hipGraphNode_t failed = nullptr;
        hipGraphExecUpdateResult result = hipGraphExecUpdateError;
        updated = hipGraphExecUpdate(root->batch_exec, g, &failed, &result) == hipSuccess &&
                  result == hipGraphExecUpdateSuccess;
        if (!updated) (void)hipGetLastError();  // a refused update leaves no error behind
      }
      updated_graph = updated;
      if (updated) {
        ++root->batch_updates;
      } else {
        if (root->batch_exec) (void)hipGraphExecDestroy(root->batch_exec);
        root->batch_exec = nullptr;
        root->batch_views.clear();
        e = hipGraphInstantiate(&root->batch_exec, g, nullptr, nullptr, 0);
        if (e == hipSuccess) {
          root->batch_views = views;
          ++root->batch_instantiations;
        } else {
          root->batch_exec = nullptr;
        }
      }
    }
    if (e == hipSuccess && status == LOTUS_HIP_PASS)
      e = hipMemcpyAsync(root->batch_vars, vars.data(), vars.size() * 4, hipMemcpyHostToDevice, root->stream);
    if (e == hipSuccess && status == LOTUS_HIP_PASS) e = hipEventRecord(start, root->stream);
    if (e == hipSuccess && status == LOTUS_HIP_PASS) e = hipGraphLaunch(root->batch_exec, root->stream);
    // the views' signatures, for replays of this graph
    for (uint32_t i = 0; i < count; ++i)
      renderers[i]->graph_signature = (e == hipSuccess && status == LOTUS_HIP_PASS && root->batch_exec)
                                          ? std::move(signatures[i])
                                          : std::vector<uint8_t>{};
  }
  if (e == hipSuccess) e = hipEventRecord(stop, root->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(root->stream);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess || status != LOTUS_HIP_PASS) {
    for (uint32_t i = 0; i < count; ++i) renderers[i]->deferred.pending = false;
    if (e != hipSuccess) set_detail(detail, detail_size, latch(root, "batched frame", e));
    return LOTUS_HIP_FAIL;
  }
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, start, stop);
  if (gpu_ms) *gpu_ms = ms;
  const uint32_t graph_kind = !graph ? 0u : updated_graph ? 2u : 1u;
  std::vector<std::vector<uint32_t>> hosts;
  e = batch_readback(renderers, params, results, count, root->stream, hosts);
  if (e != hipSuccess) {
    for (uint32_t i = 0; i < count; ++i) renderers[i]->deferred.pending = false;
    set_detail(detail, detail_size, latch(root, "batch readback", e));
    return LOTUS_HIP_FAIL;
  }
  for (uint32_t i = 0; i < count; ++i) {
    (void)batch_finish(renderers[i], &params[i], results ? &results[i] : nullptr, ms, hosts[i]);
    if (results) results[i].batch_graph = graph_kind;
  }
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

int lotus_hip_autotune(lotus_hip_renderer* r, const lotus_hip_frame_params* params, uint32_t trials,
                       lotus_hip_tuning* chosen, char* detail, size_t detail_size) {
  if (!r || !params) return LOTUS_HIP_FAIL;
  std::vector<lotus_hip_tuning> candidates;
  for (uint32_t pct : {100u, 50u}) {
    for (uint32_t div : {16u, 0u})
      for (uint32_t cam : {0u, 1u})
        for (uint32_t tr : {0u, 1u}) candidates.push_back(lotus_hip_tuning{cam, tr, div, pct, 0u, 0u});
    candidates.push_back(lotus_hip_tuning{0u, 0u, 16u, pct, 1u, 0u});  // the path megakernel
  }
  double best = 1e300;
  lotus_hip_tuning pick = r->tuning;
  lotus_hip_frame_params P = *params;
  P.frame_count = 1;
  P.max_samples = 0;
  trials = std::max(trials, 1u);
  int status = LOTUS_HIP_PASS;
  const auto trial = [&](const lotus_hip_tuning& c) {
    r->tuning = c;
    double total = 0.0;
    for (uint32_t k = 0; k <= trials && status == LOTUS_HIP_PASS; ++k) {
      lotus_hip_frame_result res{};
      r->key_valid = false;  // trial frames never join an accumulation
      status = lotus_hip_render(r, &P, &res, detail, detail_size);
      if (k > 0) total += res.frame_gpu_ms;  // the first frame sizes the rounds
    }
    if (status == LOTUS_HIP_PASS && total < best) {
      best = total;
      pick = c;
    }
  };
  for (const auto& c : candidates) {
    trial(c);
    if (status != LOTUS_HIP_PASS) return status;
  }
  // then the camera kernel's tile shapes (the block shape of its waves) for
  // the chosen wavefront tuning
  if (pick.megakernel == 0) {
    const lotus_hip_tuning base = pick;
    for (uint32_t tile : {1u, 2u}) {
      lotus_hip_tuning c = base;
      c.camera_tile = tile;
      trial(c);
      if (status != LOTUS_HIP_PASS) return status;
    }
  }
  r->tuning = pick;
  r->key_valid = false;
  if (chosen) *chosen = pick;
  set_detail(detail, detail_size, "");
  return LOTUS_HIP_PASS;
}

}  // extern "C"
