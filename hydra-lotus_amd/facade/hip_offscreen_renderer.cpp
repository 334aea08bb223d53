// Lotus::OffscreenRenderer over liblotus_hip (see hip_offscreen_renderer.hpp).
// Each method cites the reference member it stands in for
// (include/lotus/vulkan_backend.hpp) and the reference code whose mapping it
// follows.
#include "hip_offscreen_renderer.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <optional>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <variant>
#include <vector>

#include "lotus_hip.h"

namespace Lotus {
namespace {

constexpr std::uint32_t kMaterialWords = 36;  // GpuMaterialRecord, gpu_scene.hpp:94-113
constexpr std::uint32_t kBindingWords = 24;   // GpuTextureInputRecord, gpu_scene.hpp:64-80
constexpr std::uint32_t kNoCorner = 0xFFFFFFFFu;
constexpr std::uint32_t kLookupTexture = 1, kLookupFallback = 2;  // gpu_scene.hpp:56-59

FrameStatus StatusOf(int status) {
  return status == LOTUS_HIP_PASS ? FrameStatus::Pass
                                  : (status == LOTUS_HIP_SKIP ? FrameStatus::Skip : FrameStatus::Fail);
}

void PutFloat(std::uint32_t* words, std::size_t at, float value) { std::memcpy(words + at, &value, 4); }

template <std::size_t N>
void PutFloats(std::uint32_t* words, std::size_t at, const std::array<float, N>& values) {
  for (std::size_t i = 0; i < N; ++i) PutFloat(words, at + i, values[i]);
}

// One material table entry as GpuScene::Apply writes it (gpu_scene.cpp:
// 2035-2118): the 36-word record, with its lookups appended to `bindings`.
void FlattenMaterial(const SceneMaterial& entry, const std::map<const Texture*, std::uint32_t>& texture_slot,
                     std::vector<std::uint32_t>& records, std::vector<std::uint32_t>& bindings) {
  const Material& material = entry.material;
  const MaterialGeometry& geometry = material.geometry;
  const std::size_t base = records.size();
  records.resize(base + kMaterialWords, 0u);
  std::uint32_t* w = records.data() + base;
  w[0] = static_cast<std::uint32_t>(material.Model());
  w[1] = static_cast<std::uint32_t>(bindings.size() / kBindingWords);
  PutFloat(w, 4, geometry.texcoord_fallback[0]);
  PutFloat(w, 5, geometry.texcoord_fallback[1]);
  PutFloat(w, 6, geometry.opacity);
  PutFloat(w, 7, geometry.opacity_threshold);
  PutFloats(w, 8, geometry.normal);
  if (const auto* gltf = std::get_if<GltfPbr>(&material.surface)) {
    PutFloats(w, 12, gltf->base_color.value);
    PutFloat(w, 15, gltf->roughness.value);
    PutFloats(w, 16, gltf->emissive.value);
    PutFloat(w, 19, gltf->metallic.value);
    PutFloat(w, 20, gltf->emissive_strength);
    // the dielectric's reflectance at normal incidence, in double
    const double root = (1.0 - double(gltf->ior)) / (1.0 + double(gltf->ior));
    PutFloat(w, 21, static_cast<float>(root * root));
    PutFloat(w, 22, gltf->specular.value);
    PutFloat(w, 23, gltf->transmission.value);
    PutFloats(w, 24, gltf->specular_color.value);
    PutFloat(w, 27, gltf->clearcoat.value);
    PutFloats(w, 28, gltf->sheen_color.value);
    PutFloat(w, 31, gltf->sheen_roughness.value);
    PutFloat(w, 32, gltf->clearcoat_roughness.value);
    PutFloat(w, 33, gltf->ior);
  } else if (const auto* preview = std::get_if<PreviewSurface>(&material.surface)) {
    // flags above the model byte: specular workflow 1, thin walled 2
    w[0] |= ((preview->use_specular_workflow ? 1u : 0u) | (preview->thin_walled ? 2u : 0u)) << 8;
    PutFloats(w, 12, preview->base_color);
    PutFloat(w, 15, preview->roughness);
    PutFloats(w, 16, preview->emission);
    PutFloat(w, 19, preview->metallic);
    PutFloats(w, 20, preview->specular_color);
    PutFloat(w, 23, preview->ior);
    PutFloats(w, 24, preview->transmission_color);
    PutFloat(w, 27, preview->transmission);
  }
  // StandardSurface and OpenPbr carry no parameters yet: the model tag alone,
  // which the renderer refuses with its reason.
  const auto inputs = TextureInputs(material);
  for (std::size_t slot = 0; slot < inputs.size(); ++slot) {
    if (!inputs[slot]->has_value()) continue;
    w[2 + slot / 32] |= 1u << (slot % 32);
    const TextureInput& source = **inputs[slot];
    std::array<std::uint32_t, kBindingWords> b{};
    const auto found = entry.textures[slot] ? texture_slot.find(entry.textures[slot]) : texture_slot.end();
    b[0] = found == texture_slot.end() ? 0u : found->second;
    b[1] = static_cast<std::uint32_t>(source.wrap_s) * 4 + static_cast<std::uint32_t>(source.wrap_t);
    b[2] = source.channel;
    b[3] = found == texture_slot.end() ? kLookupFallback : kLookupTexture;
    PutFloats(b.data(), 4, source.scale);
    PutFloats(b.data(), 8, source.bias);
    PutFloats(b.data(), 12, source.fallback);
    for (int i = 0; i < 3; ++i) {
      PutFloat(b.data(), 16 + i, source.st_transform[i]);
      PutFloat(b.data(), 20 + i, source.st_transform[3 + i]);
    }
    bindings.insert(bindings.end(), b.begin(), b.end());
  }
}

// A Light as lotus_hip_light (render_world.hpp:89-114).
lotus_hip_light FlattenLight(const Light& light) {
  lotus_hip_light out{};
  out.shape = static_cast<std::uint32_t>(light.shape);
  out.normalize = light.normalize ? 1u : 0u;
  std::memcpy(out.world_from_light, light.world_from_light.data(), sizeof out.world_from_light);
  std::memcpy(out.radiance, light.radiance.data(), sizeof out.radiance);
  out.radius = light.radius;
  out.width = light.width;
  out.height = light.height;
  out.angle = light.angle;
  return out;
}

// previous_world_from_object * inverse(current_world_from_object), in double:
// the map of a point of this frame's placement onto the previous one's.
std::array<float, 16> PreviousFromCurrent(const Matrix4& previous, const Matrix4& current) {
  double m[16], inv[16];
  for (int i = 0; i < 16; ++i) m[i] = current[i];
  // cofactor inverse of a column-major 4x4
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  const double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  std::array<float, 16> out{};
  if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
    out[0] = out[5] = out[10] = out[15] = 1.0f;  // a singular placement maps nothing anywhere else
    return out;
  }
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += double(previous[k * 4 + r]) * (inv[c * 4 + k] / det);
      out[c * 4 + r] = static_cast<float>(s);
    }
  return out;
}

class HipOffscreenRenderer final : public OffscreenRenderer {
 public:
  HipOffscreenRenderer(lotus_hip_renderer* renderer, std::string device) : r_(renderer), device_(std::move(device)) {}
  ~HipOffscreenRenderer() override { lotus_hip_destroy(r_); }

  // vulkan_backend.hpp:760-761: the bootstrap raster pass has no HIP
  // counterpart; the scene passes are RenderScene and RenderTemporalScene.
  GpuFrameEvidence Render(const DrawSummary&, const OffscreenTarget&, std::uint32_t) override {
    GpuFrameEvidence ev;
    ev.status = FrameStatus::Skip;
    ev.detail = "the HIP renderer has no bootstrap raster pass";
    return ev;
  }

  // vulkan_backend.hpp:767: the software traversal stands in for ray queries.
  BackendCapability RayQueryCapability() const override {
    BackendCapability c;
    c.available = true;
    c.detail = "software BVH traversal on " + device_;
    return c;
  }

  // vulkan_backend.hpp:789-791.
  GpuFrameEvidence RenderScene(const DrawSummary& draw, const OffscreenTarget& target, std::uint32_t frame_count,
                               const PathTracingSettings& settings) override {
    GpuFrameEvidence ev;
    lotus_hip_frame_params p{};
    if (!FrameParams(draw.world_to_clip, target, frame_count, settings, p, ev)) return ev;
    Products products(target, settings);
    lotus_hip_frame_result res = products.Result();
    char detail[1024] = {};
    const int status = lotus_hip_render(r_, &p, &res, detail, sizeof detail);
    Finish(status, detail, target, res, products, ev);
    return ev;
  }

  // vulkan_backend.hpp:807-810. The snapshot's SceneExtraction update must
  // have been applied first (vulkan_backend.cpp:1719-1721); the instance
  // identities follow the reference's rules (temporal_metadata.cpp:6-63):
  // a placement keeps its id while FindTemporalPlacement pairs it with the
  // previous frame's, a material keeps its id while its binding is equal.
  GpuFrameEvidence RenderTemporalScene(const FrameSnapshot& snapshot, const OffscreenTarget& target,
                                       std::uint32_t frame_count, const TemporalFrameSettings& temporal,
                                       const PathTracingSettings& settings) override {
    GpuFrameEvidence ev;
    if (snapshot.scene != uploaded_scene_) {
      ev.status = FrameStatus::Fail;
      ev.detail = "apply this snapshot's SceneExtraction update before temporal rendering";
      return ev;
    }
    lotus_hip_frame_params p{};
    if (!FrameParams(ExtractDrawSummary(snapshot).world_to_clip, target, frame_count, settings, p, ev)) return ev;
    lotus_hip_temporal_settings ts{};
    ts.frame_index = temporal.frame_index;
    ts.history_key = temporal.history_key;
    ts.camera_cut = temporal.camera_cut ? 1u : 0u;
    ts.accumulate_radiance = temporal.accumulate_radiance ? 1u : 0u;
    ts.max_history_samples = temporal.max_history_samples;
    ts.filter_iterations = temporal.filter_iterations;
    ts.settle = temporal.settle ? 1u : 0u;
    ts.history_rectification = static_cast<std::uint32_t>(temporal.history_rectification);
    std::memcpy(ts.view, snapshot.camera.view.data(), sizeof ts.view);
    ts.validation[0] = temporal.validation.absolute_depth_tolerance;
    ts.validation[1] = temporal.validation.relative_depth_tolerance;
    ts.validation[2] = temporal.validation.minimum_normal_cosine;
    ts.validation[3] = temporal.validation.max_motion_pixels;
    ts.restir_temporal = settings.restir_temporal ? 1u : 0u;
    ts.restir_spatial = settings.restir_spatial ? 1u : 0u;
    ts.restir_spatial_radius = settings.restir_spatial_radius;
    ts.restir_max_history_candidates = settings.restir_max_history_candidates;
    ts.restir_spatial_neighbours = LOTUS_HIP_SPATIAL_AXES;

    // TemporalHistory::Advance decides which previous frame the identities
    // pair with (temporal.cpp:114-147); the library applies the same rules to
    // its own history and reports the reset reason.
    TemporalHistory next_history = history_;
    TemporalFrameSettings frame_settings = temporal;
    frame_settings.width = target.width;
    frame_settings.height = target.height;
    frame_settings.display_window = target.display_window;
    frame_settings.data_window = target.data_window;
    TemporalFrame frame;
    try {
      frame = next_history.Advance(snapshot, frame_settings);
    } catch (const std::invalid_argument& error) {
      ev.status = FrameStatus::Fail;
      ev.detail = error.what();
      return ev;
    }
    std::vector<std::uint64_t> instance_ids, material_ids;
    std::vector<std::uint32_t> flags;
    std::vector<float> previous_from_current;
    Identities next = identities_;
    next.Build(frame, instance_ids, material_ids, flags, previous_from_current);
    if (instance_ids.size() != instance_count_) {
      ev.status = FrameStatus::Fail;
      ev.detail = "temporal instance order/scene mismatch";
      return ev;
    }
    ts.instance_ids = instance_ids.data();
    ts.material_ids = material_ids.data();
    ts.instance_flags = flags.data();
    ts.previous_from_current = previous_from_current.data();

    Products products(target, settings);
    lotus_hip_frame_result res = products.Result();
    const std::size_t pixels = std::size_t(target.width) * target.height;
    ev.temporal.current.resize(pixels);
    ev.temporal.validity.resize(pixels);
    std::vector<TemporalRadiancePixel> radiance(temporal.accumulate_radiance ? pixels : 0);
    ts.surface = ev.temporal.current.data();
    ts.validity = ev.temporal.validity.data();
    ts.radiance = radiance.empty() ? nullptr : radiance.data();
    char detail[1024] = {};
    const int status = lotus_hip_render_temporal(r_, &p, &ts, &res, detail, sizeof detail);
    Finish(status, detail, target, res, products, ev);
    if (ev.status != FrameStatus::Pass) {
      ev.temporal = {};
      return ev;
    }
    // only a successful frame advances the stream (temporal.hpp:108-111)
    identities_ = std::move(next);
    history_ = std::move(next_history);
    TemporalProducts& t = ev.temporal;
    t.available = true;
    t.width = target.width;
    t.height = target.height;
    t.row_pitch = target.width * sizeof(TemporalPixel);
    t.validity_row_pitch = target.width * sizeof(TemporalValidityPixel);
    t.reset_reason = static_cast<TemporalResetReason>(ts.reset_reason);
    t.history_available = ts.history_available != 0;
    t.timestamp_available = true;
    t.primary_surface_gpu_ms = res.temporal_surface_gpu_ms;
    t.radiance_available = !radiance.empty();
    t.radiance = std::move(radiance);
    t.radiance_row_pitch = target.width * sizeof(TemporalRadiancePixel);
    t.batch_sample_index = ts.batch_sample_index;
    t.batch_samples = ts.batch_samples;
    t.filter_iterations = temporal.accumulate_radiance && !ts.settling ? temporal.filter_iterations : 0u;
    t.filter_gpu_ms = res.svgf_gpu_ms;
    t.settling = ts.settling != 0;
    t.settled_samples = ts.settled_samples;
    return ev;
  }

  // vulkan_backend.hpp:816-817, applied as GpuScene::Apply applies a plan
  // (gpu_scene.cpp:1606-2450): a replacement takes its previous geometry's
  // slot (gpu_scene.cpp:1651-1671), then the other releases, then the other
  // uploads. The plan vouches for what the update leaves unchanged
  // (lotus_hip_scene_plan.unchanged), so the library compares only the rest.
  GpuSceneEvidence UpdateScene(const SceneUpdate& update) override {
    GpuSceneEvidence ev;
    std::unordered_set<const MeshGeometry*> replaced_sources, replaced_targets;
    std::vector<std::uint32_t> replaced_slots;
    bool same_topology = !update.instances_changed && !update.materials_changed && update.texture_uploads.empty() &&
                         update.texture_releases.empty() && has_plan_;
    for (const auto& rep : update.geometry_replacements) {
      const auto held = std::find_if(geometries_.begin(), geometries_.end(),
                                     [&](const auto& g) { return g.get() == rep.previous; });
      const auto next = std::find_if(update.geometry_uploads.begin(), update.geometry_uploads.end(),
                                     [&](const auto& g) { return g.get() == rep.replacement; });
      if (held == geometries_.end() || next == update.geometry_uploads.end()) continue;
      const MeshGeometry& a = **held;
      const MeshGeometry& b = **next;
      same_topology = same_topology && a.positions.size() == b.positions.size() && a.triangles == b.triangles &&
                      a.normals == b.normals && a.texcoords == b.texcoords && a.colors == b.colors;
      replaced_slots.push_back(static_cast<std::uint32_t>(held - geometries_.begin()));
      *held = *next;
      replaced_sources.insert(rep.previous);
      replaced_targets.insert(rep.replacement);
    }
    for (const MeshGeometry* g : update.geometry_releases) {
      if (replaced_sources.count(g)) continue;
      same_topology = false;
      geometries_.erase(std::remove_if(geometries_.begin(), geometries_.end(),
                                       [g](const auto& held) { return held.get() == g; }),
                        geometries_.end());
    }
    for (const auto& g : update.geometry_uploads) {
      if (replaced_targets.count(g.get())) continue;
      same_topology = false;
      geometries_.push_back(g);
    }
    for (const Texture* t : update.texture_releases) {
      textures_.erase(std::remove_if(textures_.begin(), textures_.end(),
                                     [t](const auto& held) { return held.get() == t; }),
                      textures_.end());
    }
    for (const auto& t : update.texture_uploads) textures_.push_back(t);
    if (update.instances_changed) instances_ = update.instances;
    if (update.materials_changed) materials_ = update.materials;
    if (update.environment_changed) {
      environment_ = update.environment;
      environment_map_ = update.environment_map;
    }
    if (update.lights_changed) lights_ = update.lights;
    uploaded_scene_ = update.source_scene;

    Plan plan;
    std::string error;
    if (!BuildPlan(plan, error)) {
      ev.status = FrameStatus::Fail;
      ev.detail = error;
      return ev;
    }
    if (same_topology) {
      plan.plan.unchanged |= LOTUS_HIP_UNCHANGED_TOPOLOGY | LOTUS_HIP_UNCHANGED_POSITIONS;
      plan.plan.moved_geometry_count = static_cast<std::uint32_t>(replaced_slots.size());
      plan.plan.moved_geometries = replaced_slots.empty() ? nullptr : replaced_slots.data();
    }
    if (has_plan_ && !update.instances_changed) plan.plan.unchanged |= LOTUS_HIP_UNCHANGED_TRANSFORMS;
    if (has_plan_ && !update.instances_changed && !update.materials_changed)
      plan.plan.unchanged |= LOTUS_HIP_UNCHANGED_MATERIALS;
    if (has_plan_ && !update.lights_changed && !update.environment_changed)
      plan.plan.unchanged |= LOTUS_HIP_UNCHANGED_LIGHTS;
    lotus_hip_scene_evidence se{};
    char detail[1024] = {};
    const int status = lotus_hip_scene_apply(r_, &plan.plan, &se, detail, sizeof detail);
    has_plan_ = status == LOTUS_HIP_PASS;
    ev.status = StatusOf(status);
    ev.detail = detail;
    instance_count_ = plan.instance_count;
    device_bytes_ = se.device_bytes;
    GpuSceneStats& s = ev.stats;
    s.source_revision = update.source_revision;
    s.resident_geometries = static_cast<std::uint32_t>(geometries_.size());
    s.instance_count = se.instance_count;
    s.material_count = se.material_count;
    s.emissive_triangles = se.light_count > plan.plan.light_count ? se.light_count - plan.plan.light_count : 0u;
    s.analytic_lights = plan.plan.light_count;
    s.environment_map = environment_map_.has_value();
    s.environment_sampled = environment_map_.has_value();
    s.resident_textures = static_cast<std::uint32_t>(textures_.size());
    s.material_writes = se.material_writes;
    s.light_writes = se.light_writes;
    s.uploaded_bytes = se.uploaded_bytes;
    s.acceleration_available = status == LOTUS_HIP_PASS;
    s.acceleration_detail = "software wide BVH";
    s.blas_count = 1;
    s.tlas_instance_count = se.instance_count;
    s.acceleration_bytes = se.device_bytes;
    s.blas_builds = se.bvh_builds;
    s.blas_updates = se.bvh_refits;
    s.timestamps_available = true;
    ev.timings.available = true;
    ev.timings.blas_build_gpu_ms = se.build_gpu_ms;
    return ev;
  }

  // vulkan_backend.hpp:821: the device's geometry, instances and material
  // records read back; per-corner attributes are the resident host copies
  // the plan uploaded.
  GpuSceneContents ReadBackScene() override {
    GpuSceneContents out;
    lotus_hip_scene_contents c{};
    char detail[1024] = {};
    int status = lotus_hip_readback_scene(r_, &c, detail, sizeof detail);
    if (status != LOTUS_HIP_PASS) {
      out.status = StatusOf(status);
      out.detail = detail;
      return out;
    }
    std::vector<float> positions(c.position_floats);
    std::vector<std::uint32_t> triangles(c.index_words), vo(c.geometry_count), vc(c.geometry_count),
        to(c.geometry_count), tc(c.geometry_count), ig(c.instance_count), im(c.instance_count),
        materials(std::size_t(c.material_count) * kMaterialWords);
    std::vector<float> xf(std::size_t(c.instance_count) * 16);
    c.positions = positions.data();
    c.triangles = triangles.data();
    c.geometry_vertex_offset = vo.data();
    c.geometry_vertex_count = vc.data();
    c.geometry_triangle_offset = to.data();
    c.geometry_triangle_count = tc.data();
    c.instance_geometry = ig.data();
    c.instance_transform = xf.data();
    c.instance_material = im.data();
    c.materials = materials.data();
    status = lotus_hip_readback_scene(r_, &c, detail, sizeof detail);
    out.status = StatusOf(status);
    out.detail = detail;
    if (status != LOTUS_HIP_PASS) return out;
    for (std::uint32_t g = 0; g < c.geometry_count; ++g) {
      GpuGeometryContents gc;
      gc.slot = g;
      for (std::uint32_t v = 0; v < vc[g]; ++v) {
        const float* p = positions.data() + 3ull * (vo[g] + v);
        gc.positions.push_back({p[0], p[1], p[2]});
      }
      for (std::uint32_t t = 0; t < tc[g]; ++t) {
        const std::uint32_t* i = triangles.data() + 3ull * (to[g] + t);
        gc.triangles.push_back({i[0], i[1], i[2]});
      }
      if (g < geometries_.size()) {
        const MeshGeometry& held = *geometries_[g];
        gc.normals = held.normals;
        for (const auto& [name, set] : held.texcoords) gc.texcoords.push_back(set);
        for (const auto& [name, set] : held.colors) gc.colors.push_back(set);
      }
      out.geometries.push_back(std::move(gc));
    }
    for (std::uint32_t i = 0; i < c.instance_count; ++i) {
      GpuInstanceContents ic;
      std::copy(xf.begin() + 16 * i, xf.begin() + 16 * (i + 1), ic.world_from_object.begin());
      ic.geometry_slot = ig[i];
      ic.material_slot = im[i];
      if (i < instances_.size()) {
        ic.texcoords = instances_[i].texcoords;
        ic.colors = instances_[i].colors;
      }
      GpuTlasInstanceContents tlas;
      for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 4; ++col) tlas.object_to_world[r * 4 + col] = ic.world_from_object[col * 4 + r];
      tlas.geometry_slot = ig[i];
      tlas.mask = 0xFF;
      out.tlas_instances.push_back(tlas);
      out.instances.push_back(ic);
    }
    for (std::uint32_t m = 0; m < c.material_count && m < materials_.size(); ++m) {
      GpuMaterialContents mc;
      mc.material = materials_[m].material;
      for (std::size_t slot = 0; slot < kMaterialTextureInputs; ++slot) {
        const Texture* t = materials_[m].textures[slot];
        for (std::uint32_t k = 0; t && k < textures_.size(); ++k)
          if (textures_[k].get() == t) mc.texture_slots[slot] = k;
      }
      out.materials.push_back(std::move(mc));
    }
    for (std::uint32_t k = 0; k < textures_.size(); ++k) out.textures.push_back({k, *textures_[k]});
    out.environment = {c.environment[0], c.environment[1], c.environment[2]};
    return out;
  }

  // vulkan_backend.hpp:835-836: presentation stays with the Vulkan window
  // layer.
  PresentEvidence Present(std::uint32_t, std::uint32_t, const PresentView&) override {
    PresentEvidence ev;
    ev.status = FrameStatus::Fail;
    ev.detail = "the HIP renderer has no presentation surface";
    return ev;
  }

  // vulkan_backend.hpp:841.
  GpuResidency Residency() const override {
    GpuResidency r;
    r.device_local_bytes = device_bytes_;
    return r;
  }

 private:
  struct Plan {
    lotus_hip_scene_plan plan{};
    std::uint32_t instance_count = 0;
    std::vector<float> positions, normals, texcoords, colors, transforms;
    std::vector<std::uint32_t> triangles, vertex_offset, vertex_count, triangle_offset, triangle_count,
        normal_offset, texcoord_offset, instance_geometry, instance_material, instance_texcoord,
        instance_color, instance_color_constant, materials, bindings;
    std::vector<lotus_hip_light> lights;
    std::vector<lotus_hip_texture> textures;
    lotus_hip_environment_map environment_map{};
  };

  // Identities of instances and materials across temporal frames
  // (temporal_metadata.cpp:6-63).
  struct Identities {
    std::optional<FrameSnapshot> last;
    std::map<std::pair<std::string, std::size_t>, std::uint64_t> placements;
    std::map<std::string, std::pair<SceneMaterial, std::uint64_t>> materials;
    std::uint64_t next_id = 1;

    void Build(const TemporalFrame& frame, std::vector<std::uint64_t>& instance_ids,
               std::vector<std::uint64_t>& material_ids, std::vector<std::uint32_t>& flags,
               std::vector<float>& previous_from_current) {
      std::map<std::pair<std::string, std::size_t>, std::uint64_t> now_placements;
      std::map<std::string, std::pair<SceneMaterial, std::uint64_t>> now_materials;
      TemporalFrame identity_frame = frame;
      identity_frame.previous = last;
      if (frame.current.scene) {
        const LotusScene& scene = *frame.current.scene;
        for (const auto& [key, mesh] : scene.meshes) {
          if (!mesh.instance.visible || !mesh.geometry || mesh.geometry->triangles.empty()) continue;
          auto found = now_materials.find(mesh.material);
          if (found == now_materials.end()) {
            SceneMaterial bound;
            if (const auto it = scene.materials.find(mesh.material); it != scene.materials.end())
              bound.material = it->second;
            const auto inputs = TextureInputs(bound.material);
            for (std::size_t i = 0; i < inputs.size(); ++i) {
              if (!inputs[i]->has_value()) continue;
              const auto t = scene.textures.find((*inputs[i])->texture);
              if (t != scene.textures.end()) bound.textures[i] = t->second.get();
            }
            const auto old = materials.find(mesh.material);
            const std::uint64_t id =
                old != materials.end() && old->second.first == bound ? old->second.second : next_id++;
            found = now_materials.emplace(mesh.material, std::make_pair(bound, id)).first;
          }
          const auto transforms = PlacementTransforms(mesh.instance);
          for (std::size_t ordinal = 0; ordinal < transforms.size(); ++ordinal) {
            const std::pair<std::string, std::size_t> pk{key, ordinal};
            const auto old = placements.find(pk);
            const bool same = old != placements.end() && FindTemporalPlacement(identity_frame, key, ordinal);
            const std::uint64_t id = same ? old->second : next_id++;
            now_placements.emplace(pk, id);
            instance_ids.push_back(id);
            material_ids.push_back(found->second.second);
            std::array<float, 16> map{};
            map[0] = map[5] = map[10] = map[15] = 1.0f;
            std::uint32_t f = 0;
            if (const auto pair = FindTemporalPlacement(frame, key, ordinal)) {
              map = PreviousFromCurrent(pair->previous_world_from_object, pair->current_world_from_object);
              f = 2u | (pair->material_unchanged ? 4u : 0u);
            }
            flags.push_back(f);
            previous_from_current.insert(previous_from_current.end(), map.begin(), map.end());
          }
        }
      }
      placements = std::move(now_placements);
      materials = std::move(now_materials);
      last = frame.current;
    }
  };

  // The target's colour and depth read back into the evidence's products
  // (vulkan_backend.cpp:1305-1326): RGBA32F colour and D32 depth, rows from
  // the top.
  struct Products {
    std::vector<float> color, depth;
    std::vector<std::uint32_t> counts;
    Products(const OffscreenTarget& target, const PathTracingSettings& settings) {
      const std::size_t pixels = std::size_t(target.width) * target.height;
      if (target.read_back) {
        color.resize(pixels * 4);
        depth.resize(pixels);
      }
      counts.resize(std::size_t(std::min(settings.max_bounces, kMaxWavefrontBounces)) + 2);
    }
    lotus_hip_frame_result Result() {
      lotus_hip_frame_result r{};
      r.color = color.empty() ? nullptr : color.data();
      r.depth = depth.empty() ? nullptr : depth.data();
      r.path_counts = counts.data();
      r.path_counts_capacity = static_cast<std::uint32_t>(counts.size());
      return r;
    }
  };

  bool FrameParams(const Matrix4& world_to_clip, const OffscreenTarget& target, std::uint32_t frame_count,
                   const PathTracingSettings& s, lotus_hip_frame_params& p, GpuFrameEvidence& ev) const {
    if (s.environment_background) {
      ev.status = FrameStatus::Fail;
      ev.detail = "the HIP renderer shows the clear colour behind missed camera samples";
      return false;
    }
    std::memcpy(p.world_to_clip, world_to_clip.data(), sizeof p.world_to_clip);
    p.width = target.width;
    p.height = target.height;
    std::memcpy(p.display_window, target.display_window.data(), sizeof p.display_window);
    std::memcpy(p.data_window, target.data_window.data(), sizeof p.data_window);
    std::memcpy(p.clear_color, target.clear_color.data(), sizeof p.clear_color);
    p.frame_count = frame_count;
    p.output = static_cast<std::uint32_t>(s.output);
    p.light_sampling = static_cast<std::uint32_t>(s.light_sampling);
    p.sample_index = s.sample_index;
    p.max_bounces = s.max_bounces;
    p.max_samples = s.max_samples;
    p.restir_candidates = s.restir_candidates;
    p.restir_proposal = static_cast<std::uint32_t>(s.restir_proposal);
    return true;
  }

  void Finish(int status, const char* detail, const OffscreenTarget& target, const lotus_hip_frame_result& res,
              Products& products, GpuFrameEvidence& ev) {
    ev.status = StatusOf(status);
    ev.detail = detail;
    ev.device_name = device_;
    if (status != LOTUS_HIP_PASS) return;
    ev.frames_rendered = res.frames_rendered;
    ev.samples_per_pixel = res.samples_per_pixel;
    ev.ray_query_used = true;
    if (target.read_back) {
      ev.color.width = ev.depth.width = target.width;
      ev.color.height = ev.depth.height = target.height;
      ev.color.row_pitch = target.width * 16;
      ev.color.pixel_format = "rgba32-sfloat";
      ev.color.origin = "top-left";
      ev.color.color_space = "linear";
      ev.color.payload.resize(products.color.size() * 4);
      std::memcpy(ev.color.payload.data(), products.color.data(), ev.color.payload.size());
      ev.depth.row_pitch = target.width * 4;
      ev.depth.pixel_format = "d32-sfloat";
      ev.depth.origin = "top-left";
      ev.depth.payload = std::move(products.depth);
    }
    products.counts.resize(res.path_counts_size);
    ev.wavefront_path_counts = std::move(products.counts);
    ev.wavefront_rounds = res.wavefront_rounds;
    WavefrontTimings& t = ev.wavefront_timings;
    t.available = res.frames_rendered != 0;
    t.camera_gpu_ms = res.camera_gpu_ms;
    t.intersect_gpu_ms = res.trace_gpu_ms + res.intersect_gpu_ms + res.shadow_gpu_ms;
    t.shade_gpu_ms = res.shade_gpu_ms;
    t.tail_gpu_ms = res.tail_gpu_ms;
  }

  // The resident scene as one plan: geometries in residency order, each
  // geometry's corner attributes after its predecessors', and per instance
  // the texture-coordinate and colour set its SceneInstance names.
  bool BuildPlan(Plan& out, std::string& error) {
    std::unordered_map<const MeshGeometry*, std::uint32_t> slot_of;
    std::vector<std::vector<std::uint32_t>> texcoord_set_first(geometries_.size()), color_set_first(geometries_.size()),
        color_set_size(geometries_.size());
    bool any_normals = false, any_texcoords = false, any_colors = false;
    for (std::uint32_t g = 0; g < geometries_.size(); ++g) {
      const MeshGeometry& m = *geometries_[g];
      slot_of[&m] = g;
      const std::uint32_t corners = static_cast<std::uint32_t>(m.triangles.size() * 3);
      out.vertex_offset.push_back(static_cast<std::uint32_t>(out.positions.size() / 3));
      out.vertex_count.push_back(static_cast<std::uint32_t>(m.positions.size()));
      out.triangle_offset.push_back(static_cast<std::uint32_t>(out.triangles.size() / 3));
      out.triangle_count.push_back(static_cast<std::uint32_t>(m.triangles.size()));
      for (const auto& p : m.positions) out.positions.insert(out.positions.end(), p.begin(), p.end());
      for (const auto& t : m.triangles) out.triangles.insert(out.triangles.end(), t.begin(), t.end());
      if (!m.normals.empty() && m.normals.size() != corners) {
        error = "a geometry's normals must have one entry per triangle corner";
        return false;
      }
      out.normal_offset.push_back(m.normals.empty() ? kNoCorner : static_cast<std::uint32_t>(out.normals.size() / 3));
      for (const auto& n : m.normals) out.normals.insert(out.normals.end(), n.begin(), n.end());
      any_normals |= !m.normals.empty();
      std::uint32_t first_set = kNoCorner;
      for (const auto& [name, set] : m.texcoords) {
        if (set.size() != corners) {
          error = "a texture-coordinate set must have one entry per triangle corner";
          return false;
        }
        const std::uint32_t first = static_cast<std::uint32_t>(out.texcoords.size() / 2);
        if (first_set == kNoCorner) first_set = first;
        texcoord_set_first[g].push_back(first);
        for (const auto& c : set) out.texcoords.insert(out.texcoords.end(), c.begin(), c.end());
        any_texcoords = true;
      }
      out.texcoord_offset.push_back(first_set);
      for (const auto& [name, set] : m.colors) {
        if (set.size() != corners && set.size() != 1) {
          error = "a colour set must have one entry per triangle corner or one for the mesh";
          return false;
        }
        color_set_first[g].push_back(static_cast<std::uint32_t>(out.colors.size() / 3));
        color_set_size[g].push_back(static_cast<std::uint32_t>(set.size()));
        for (const auto& c : set) out.colors.insert(out.colors.end(), c.begin(), c.end());
        any_colors = true;
      }
    }
    for (const SceneInstance& i : instances_) {
      const auto found = slot_of.find(i.geometry);
      if (found == slot_of.end()) {
        error = "an instance places a geometry that is not resident";
        return false;
      }
      const std::uint32_t g = found->second;
      out.instance_geometry.push_back(g);
      out.transforms.insert(out.transforms.end(), i.world_from_object.begin(), i.world_from_object.end());
      out.instance_material.push_back(i.material);
      out.instance_texcoord.push_back(i.texcoords < texcoord_set_first[g].size() ? texcoord_set_first[g][i.texcoords]
                                                                                  : kNoCorner);
      const bool has_colors = i.colors < color_set_first[g].size();
      out.instance_color.push_back(has_colors ? color_set_first[g][i.colors] : kNoCorner);
      out.instance_color_constant.push_back(has_colors && color_set_size[g][i.colors] == 1 ? 1u : 0u);
    }
    std::map<const Texture*, std::uint32_t> texture_slot;
    for (std::uint32_t k = 0; k < textures_.size(); ++k) {
      const Texture& t = *textures_[k];
      texture_slot[&t] = k;
      out.textures.push_back(lotus_hip_texture{t.width, t.height, static_cast<std::uint32_t>(t.format),
                                               t.texels.data()});
    }
    if (materials_.empty()) materials_.push_back(SceneMaterial{});  // the default material at slot 0
    for (const SceneMaterial& m : materials_) FlattenMaterial(m, texture_slot, out.materials, out.bindings);
    for (const Light& l : lights_) out.lights.push_back(FlattenLight(l));

    lotus_hip_scene_plan& p = out.plan;
    p.geometry_count = static_cast<std::uint32_t>(geometries_.size());
    p.positions = out.positions.data();
    p.triangles = out.triangles.data();
    p.geometry_vertex_offset = out.vertex_offset.data();
    p.geometry_vertex_count = out.vertex_count.data();
    p.geometry_triangle_offset = out.triangle_offset.data();
    p.geometry_triangle_count = out.triangle_count.data();
    p.instance_count = out.instance_count = static_cast<std::uint32_t>(instances_.size());
    p.instance_geometry = out.instance_geometry.data();
    p.instance_transform = out.transforms.data();
    p.instance_material = out.instance_material.data();
    p.material_count = static_cast<std::uint32_t>(materials_.size());
    p.materials = out.materials.data();
    std::memcpy(p.environment, environment_.data(), sizeof p.environment);
    p.light_count = static_cast<std::uint32_t>(out.lights.size());
    p.lights = out.lights.empty() ? nullptr : out.lights.data();
    if (environment_map_ && environment_map_->texture) {
      const Texture& t = *environment_map_->texture;
      out.environment_map.width = t.width;
      out.environment_map.height = t.height;
      out.environment_map.texels = t.texels.data();
      // TextureFormat's values are LOTUS_HIP_TEXTURE_*'s, as for the bindings' textures
      out.environment_map.format = static_cast<std::uint32_t>(t.format);
      std::memcpy(out.environment_map.scale, environment_map_->scale.data(), sizeof out.environment_map.scale);
      std::memcpy(out.environment_map.world_from_environment, environment_map_->world_from_environment.data(),
                  sizeof out.environment_map.world_from_environment);
      p.environment_map = &out.environment_map;
    }
    p.texture_count = static_cast<std::uint32_t>(out.textures.size());
    p.textures = out.textures.empty() ? nullptr : out.textures.data();
    p.binding_count = static_cast<std::uint32_t>(out.bindings.size() / kBindingWords);
    p.bindings = out.bindings.empty() ? nullptr : out.bindings.data();
    p.texcoords = any_texcoords ? out.texcoords.data() : nullptr;
    p.geometry_texcoord_offset = any_texcoords ? out.texcoord_offset.data() : nullptr;
    p.instance_texcoord_offset = any_texcoords ? out.instance_texcoord.data() : nullptr;
    p.normals = any_normals ? out.normals.data() : nullptr;
    p.geometry_normal_offset = any_normals ? out.normal_offset.data() : nullptr;
    p.colors = any_colors ? out.colors.data() : nullptr;
    p.instance_color_offset = any_colors ? out.instance_color.data() : nullptr;
    p.instance_color_constant = any_colors ? out.instance_color_constant.data() : nullptr;
    p.bvh_builder = LOTUS_HIP_BVH_SAH;
    return true;
  }

  lotus_hip_renderer* r_;
  std::string device_;
  std::vector<std::shared_ptr<const MeshGeometry>> geometries_;
  std::vector<std::shared_ptr<const Texture>> textures_;
  std::vector<SceneInstance> instances_;
  std::vector<SceneMaterial> materials_;
  std::array<float, 3> environment_{};
  bool has_plan_ = false;  // the last plan applied: the next may vouch for parts of it
  std::optional<SceneEnvironmentMap> environment_map_;
  std::vector<Light> lights_;
  std::shared_ptr<const LotusScene> uploaded_scene_;
  std::uint32_t instance_count_ = 0;
  std::uint64_t device_bytes_ = 0;
  TemporalHistory history_;
  Identities identities_;
};

}  // namespace

std::unique_ptr<OffscreenRenderer> CreateHipOffscreenRenderer(FrameStatus& status, std::string& error, int device) {
  char detail[1024] = {};
  int st = LOTUS_HIP_FAIL;
  lotus_hip_renderer* r = lotus_hip_create(device, &st, detail, sizeof detail);
  status = StatusOf(st);
  error = detail;
  if (!r) return nullptr;
  return std::make_unique<HipOffscreenRenderer>(r, detail);
}

}  // namespace Lotus
