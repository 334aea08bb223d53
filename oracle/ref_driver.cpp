// C driver over the reference's own std-only CPU functions, compiled in place
// from the reference sources by `make -C oracle ref` into oracle/_ref/ (never
// committed). TEST INFRASTRUCTURE ONLY: tests use it to pin the oracle's
// restatements (fixtures, DES-Q5, reservoir arithmetic) against the
// reference's code where that code builds without Vulkan.
#include <cstdint>
#include <cstring>
#include <exception>
#include <map>
#include <string>
#include <string_view>
#include <vector>

#include <lotus/extraction.hpp>
#include <lotus/render_world.hpp>
#include <lotus/temporal.hpp>

#include "material_fixtures.hpp"
#include "reference.hpp"
#include "reservoir.hpp"

namespace {
void CopyString(const std::string& s, char* out, size_t size) {
  if (!out || size == 0) return;
  std::strncpy(out, s.c_str(), size - 1);
  out[size - 1] = 0;
}
}  // namespace

extern "C" {

// adapters/common/reference.cpp:286 (Compare). Images: 3 floats per pixel,
// rows from the top.
int ref_compare(const float* image, uint32_t width, uint32_t height,
                uint32_t samples, const float* mean, const float* variance,
                uint32_t reference_samples, double* image_z, double* max_tile_z,
                double* mean_tile_z2, uint32_t* tile, char* failure,
                size_t failure_size) {
  LotusHeadless::RgbImage img{width, height,
                              std::vector<float>(image, image + size_t(width) * height * 3)};
  LotusHeadless::Reference ref;
  ref.mean = {width, height, std::vector<float>(mean, mean + size_t(width) * height * 3)};
  ref.variance = {width, height,
                  std::vector<float>(variance, variance + size_t(width) * height * 3)};
  ref.samples = reference_samples;
  const auto c = LotusHeadless::Compare(img, samples, ref);
  for (int i = 0; i < 3; ++i) image_z[i] = c.image_z[i];
  *max_tile_z = c.max_tile_z;
  *mean_tile_z2 = c.mean_tile_z2;
  *tile = c.tile;
  CopyString(c.failure, failure, failure_size);
  return c.failure.empty() ? 0 : 1;
}

// adapters/common/reference.cpp:104-130 (ReadPfm): 3 floats per pixel, rows
// from the top. Call with values == nullptr for the size.
int ref_read_pfm(const char* path, uint32_t* width, uint32_t* height,
                 float* values) {
  LotusHeadless::RgbImage img;
  if (!LotusHeadless::ReadPfm(path, img).empty()) return 1;
  *width = img.width;
  *height = img.height;
  if (values) std::memcpy(values, img.values.data(), img.values.size() * 4);
  return 0;
}

// adapters/common/reference.cpp:221-280 (SetCornellBox), committed through
// core/render-world/render_world.cpp. Writes, per mesh in key order: vertex
// count, triangle count, then positions (3 floats each), triangles (3 uint
// each), the placement transform (16 floats), the material's base colour,
// roughness, metallic, ior, emission (9 floats). Also the camera's view and
// projection (32 floats). Returns the number of floats/words written, or the
// needed size when `out` is null.
size_t ref_cornell_scene(float red_scale, uint32_t* out, size_t capacity,
                         float* camera) {
  Lotus::RenderWorld world;
  LotusHeadless::SetCornellBox(world);
  if (red_scale != 1.0f) LotusHeadless::SetCornellRedWall(world, red_scale);
  const Lotus::FrameSnapshot snap = world.Commit();
  std::vector<uint32_t> words;
  auto push_f = [&](float f) {
    uint32_t w;
    std::memcpy(&w, &f, 4);
    words.push_back(w);
  };
  for (const auto& [key, mesh] : snap.scene->meshes) {
    const auto& g = *mesh.geometry;
    words.push_back(static_cast<uint32_t>(g.positions.size()));
    words.push_back(static_cast<uint32_t>(g.triangles.size()));
    for (const auto& p : g.positions)
      for (float v : p) push_f(v);
    for (const auto& t : g.triangles)
      for (uint32_t v : t) words.push_back(v);
    for (float v : mesh.instance.world_from_object) push_f(v);
    Lotus::Material m;
    const auto it = snap.scene->materials.find(mesh.material);
    if (it != snap.scene->materials.end()) m = it->second;
    const auto& p = Lotus::Preview(m);
    for (float v : p.base_color) push_f(v);
    push_f(p.roughness);
    push_f(p.metallic);
    push_f(p.ior);
    for (float v : p.emission) push_f(v);
  }
  if (camera) {
    std::memcpy(camera, snap.camera.view.data(), 64);
    std::memcpy(camera + 16, snap.camera.projection.data(), 64);
  }
  if (out && capacity >= words.size())
    std::memcpy(out, words.data(), words.size() * 4);
  return words.size();
}

// A material fixture (adapters/common/material_fixtures.cpp:2402-2435,
// MaterialFixtures) as the reference's SceneExtraction turns it into a
// SceneUpdate (core/render-extraction/extraction.cpp), the plan its
// GpuScene::Apply receives, written as a word stream of this driver's own
// layout (floats as their bits):
//   size, camera view[16], projection[16], environment[3];
//   lights: count, then per light shape, normalize, world_from_light[16],
//     radiance[3], radius, width, height, angle;
//   textures: count, then per texture width, height, format, byte count and
//     the texel bytes padded to whole words;
//   geometries: count, then per geometry vertex count, triangle count,
//     positions, triangles, normal count and normals (3 per corner),
//     texture-coordinate sets (count, then per set its length and 2 floats
//     per entry), colour sets (count, then per set its length and 3 floats
//     per entry);
//   materials: count, then per material model, the PreviewSurface fields
//     (base colour[3], roughness, metallic, ior, specular workflow,
//     specular colour[3], emission[3], transmission, transmission
//     colour[3], thin walled), the MaterialGeometry fields (texcoord
//     fallback[2], opacity, opacity threshold, normal[3]) and 14 texture
//     input slots (present, texture index, channel, wrap s, wrap t,
//     scale[4], bias[4], fallback[4], st transform[6]);
//   instances: count, then per instance geometry index,
//     world_from_object[16], material slot, texture-coordinate set, colour
//     set (0xFFFFFFFF for none).
// `control` applies the fixture's negative control first. Returns the word
// count (the needed size when `out` is null or too small), 0 for an unknown
// fixture.
size_t ref_material_fixture(const char* id, int control, uint32_t* out, size_t capacity) {
  const LotusHeadless::MaterialFixture* fixture = nullptr;
  for (const auto& f : LotusHeadless::MaterialFixtures())
    if (f.id == std::string_view(id)) fixture = &f;
  if (!fixture) return 0;
  Lotus::RenderWorld world;
  fixture->set(world);
  if (control) fixture->apply_control(world);
  const Lotus::FrameSnapshot snap = world.Commit();
  Lotus::SceneExtraction extraction;
  const Lotus::SceneUpdate update = extraction.Update(snap);
  std::vector<uint32_t> w;
  auto push_f = [&](float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    w.push_back(b);
  };
  auto push_fs = [&](const float* f, size_t n) {
    for (size_t i = 0; i < n; ++i) push_f(f[i]);
  };
  w.push_back(fixture->size);
  push_fs(snap.camera.view.data(), 16);
  push_fs(snap.camera.projection.data(), 16);
  push_fs(update.environment.data(), 3);
  w.push_back(static_cast<uint32_t>(update.lights.size()));
  for (const Lotus::Light& l : update.lights) {
    w.push_back(static_cast<uint32_t>(l.shape));
    w.push_back(l.normalize ? 1u : 0u);
    push_fs(l.world_from_light.data(), 16);
    push_fs(l.radiance.data(), 3);
    push_f(l.radius);
    push_f(l.width);
    push_f(l.height);
    push_f(l.angle);
  }
  std::map<const Lotus::Texture*, uint32_t> texture_index;
  w.push_back(static_cast<uint32_t>(update.texture_uploads.size()));
  for (const auto& t : update.texture_uploads) {
    texture_index[t.get()] = static_cast<uint32_t>(texture_index.size());
    w.push_back(t->width);
    w.push_back(t->height);
    w.push_back(static_cast<uint32_t>(t->format));
    w.push_back(static_cast<uint32_t>(t->texels.size()));
    const size_t words = (t->texels.size() + 3) / 4;
    const size_t at = w.size();
    w.resize(at + words, 0u);
    std::memcpy(w.data() + at, t->texels.data(), t->texels.size());
  }
  std::map<const Lotus::MeshGeometry*, uint32_t> geometry_index;
  w.push_back(static_cast<uint32_t>(update.geometry_uploads.size()));
  for (const auto& g : update.geometry_uploads) {
    geometry_index[g.get()] = static_cast<uint32_t>(geometry_index.size());
    w.push_back(static_cast<uint32_t>(g->positions.size()));
    w.push_back(static_cast<uint32_t>(g->triangles.size()));
    for (const auto& p : g->positions) push_fs(p.data(), 3);
    for (const auto& t : g->triangles)
      for (uint32_t v : t) w.push_back(v);
    w.push_back(static_cast<uint32_t>(g->normals.size()));
    for (const auto& n : g->normals) push_fs(n.data(), 3);
    w.push_back(static_cast<uint32_t>(g->texcoords.size()));
    for (const auto& [name, set] : g->texcoords) {
      w.push_back(static_cast<uint32_t>(set.size()));
      for (const auto& c : set) push_fs(c.data(), 2);
    }
    w.push_back(static_cast<uint32_t>(g->colors.size()));
    for (const auto& [name, set] : g->colors) {
      w.push_back(static_cast<uint32_t>(set.size()));
      for (const auto& c : set) push_fs(c.data(), 3);
    }
  }
  w.push_back(static_cast<uint32_t>(update.materials.size()));
  for (const Lotus::SceneMaterial& entry : update.materials) {
    const Lotus::Material& m = entry.material;
    w.push_back(static_cast<uint32_t>(m.Model()));
    Lotus::PreviewSurface p;
    if (const auto* preview = std::get_if<Lotus::PreviewSurface>(&m.surface)) p = *preview;
    push_fs(p.base_color.data(), 3);
    push_f(p.roughness);
    push_f(p.metallic);
    push_f(p.ior);
    w.push_back(p.use_specular_workflow ? 1u : 0u);
    push_fs(p.specular_color.data(), 3);
    push_fs(p.emission.data(), 3);
    push_f(p.transmission);
    push_fs(p.transmission_color.data(), 3);
    w.push_back(p.thin_walled ? 1u : 0u);
    push_fs(m.geometry.texcoord_fallback.data(), 2);
    push_f(m.geometry.opacity);
    push_f(m.geometry.opacity_threshold);
    push_fs(m.geometry.normal.data(), 3);
    // glTF PBR's parameters (material.hpp:180-219), defaults for other models
    Lotus::GltfPbr g;
    if (const auto* gltf = std::get_if<Lotus::GltfPbr>(&m.surface)) g = *gltf;
    push_fs(g.base_color.value.data(), 3);
    push_f(g.roughness.value);
    push_f(g.metallic.value);
    push_fs(g.emissive.value.data(), 3);
    push_f(g.emissive_strength);
    push_f(g.ior);
    push_f(g.specular.value);
    push_fs(g.specular_color.value.data(), 3);
    push_f(g.transmission.value);
    push_f(g.clearcoat.value);
    push_f(g.clearcoat_roughness.value);
    push_fs(g.sheen_color.value.data(), 3);
    push_f(g.sheen_roughness.value);
    const auto inputs = Lotus::TextureInputs(m);
    for (size_t slot = 0; slot < inputs.size(); ++slot) {
      const auto& input = *inputs[slot];
      w.push_back(input.has_value() ? 1u : 0u);
      if (!input.has_value()) continue;
      const auto found = texture_index.find(entry.textures[slot]);
      w.push_back(found == texture_index.end() ? 0xFFFFFFFFu : found->second);
      w.push_back(input->channel);
      w.push_back(static_cast<uint32_t>(input->wrap_s));
      w.push_back(static_cast<uint32_t>(input->wrap_t));
      push_fs(input->scale.data(), 4);
      push_fs(input->bias.data(), 4);
      push_fs(input->fallback.data(), 4);
      push_fs(input->st_transform.data(), 6);
    }
  }
  w.push_back(static_cast<uint32_t>(update.instances.size()));
  for (const Lotus::SceneInstance& i : update.instances) {
    w.push_back(geometry_index.at(i.geometry));
    push_fs(i.world_from_object.data(), 16);
    w.push_back(i.material);
    w.push_back(i.texcoords);
    w.push_back(i.colors);
  }
  if (out && capacity >= w.size()) std::memcpy(out, w.data(), w.size() * 4);
  return w.size();
}

// backend/vulkan/reservoir.cpp: StreamReservoirCandidate. Reservoir as 4
// words {sample, weight_sum, selected_target, M}. Returns 1 selected, 0 not,
// -1 on a thrown failure (reservoir unchanged).
int ref_stream_candidate(uint32_t* reservoir, uint32_t sample, float target,
                         float proposal_pdf, float random) {
  Lotus::vulkan_internal::Reservoir r;
  std::memcpy(&r, reservoir, 16);
  try {
    const bool s = Lotus::vulkan_internal::StreamReservoirCandidate(
        r, sample, target, proposal_pdf, random);
    std::memcpy(reservoir, &r, 16);
    return s ? 1 : 0;
  } catch (const std::exception&) {
    return -1;
  }
}

int ref_merge_reevaluated(uint32_t* destination, const uint32_t* source,
                          double source_normalization, float receiver_target,
                          uint32_t max_source_candidates, float random) {
  Lotus::vulkan_internal::Reservoir d, s;
  std::memcpy(&d, destination, 16);
  std::memcpy(&s, source, 16);
  try {
    const bool sel = Lotus::vulkan_internal::MergeReevaluatedReservoir(
        d, s, source_normalization, receiver_target, max_source_candidates, random);
    std::memcpy(destination, &d, 16);
    return sel ? 1 : 0;
  } catch (const std::exception&) {
    return -1;
  }
}

double ref_reservoir_normalization(const uint32_t* reservoir) {
  Lotus::vulkan_internal::Reservoir r;
  std::memcpy(&r, reservoir, 16);
  try {
    return Lotus::vulkan_internal::ReservoirNormalization(r);
  } catch (const std::exception&) {
    return -1.0;
  }
}

double ref_reuse_normalization(const uint32_t* reservoir, uint32_t supporting) {
  Lotus::vulkan_internal::Reservoir r;
  std::memcpy(&r, reservoir, 16);
  try {
    return Lotus::vulkan_internal::ReservoirReuseNormalization(r, supporting);
  } catch (const std::exception&) {
    return -1.0;
  }
}

// backend/vulkan/temporal.cpp:375-421 (EstimateTemporalVariance) over
// width*height TemporalPixel (64 B) and TemporalRadiancePixel (32 B) records.
float ref_estimate_variance(const void* surface, const void* radiance, uint32_t x, uint32_t y, uint32_t width,
                            uint32_t height, uint32_t batch_samples, uint32_t max_history_samples) {
  const size_t n = size_t(width) * height;
  std::span<const Lotus::TemporalPixel> s(static_cast<const Lotus::TemporalPixel*>(surface), n);
  std::span<const Lotus::TemporalRadiancePixel> r(static_cast<const Lotus::TemporalRadiancePixel*>(radiance), n);
  return Lotus::EstimateTemporalVariance(s, r, x, y, width, height, batch_samples, max_history_samples);
}

// temporal.cpp:424-512 (FilterTemporalRadiance): RGBA per pixel into out.
// Returns 0 when the oracle rejects the input.
int ref_filter_radiance(const void* surface, const void* radiance, uint32_t width, uint32_t height,
                        uint32_t batch_samples, uint32_t max_history_samples, uint32_t iterations, float* out) {
  const size_t n = size_t(width) * height;
  std::span<const Lotus::TemporalPixel> s(static_cast<const Lotus::TemporalPixel*>(surface), n);
  std::span<const Lotus::TemporalRadiancePixel> r(static_cast<const Lotus::TemporalRadiancePixel*>(radiance), n);
  const auto v = Lotus::FilterTemporalRadiance(s, r, width, height, batch_samples, max_history_samples, iterations);
  if (v.size() != n) return 0;
  std::memcpy(out, v.data(), n * 16);
  return 1;
}

// temporal.cpp:258-300 (ReprojectTemporalHistory): the validity record of
// pixel (x, y) against width*height previous records; settings: absolute and
// relative depth tolerance, minimum normal cosine, max motion pixels.
void ref_reproject(const void* current, uint32_t x, uint32_t y, uint32_t width, uint32_t height,
                   const void* previous, const float* settings, uint32_t* out) {
  Lotus::TemporalValidationSettings v;
  v.absolute_depth_tolerance = settings[0];
  v.relative_depth_tolerance = settings[1];
  v.minimum_normal_cosine = settings[2];
  v.max_motion_pixels = settings[3];
  const size_t n = previous ? size_t(width) * height : 0;
  std::span<const Lotus::TemporalPixel> p(static_cast<const Lotus::TemporalPixel*>(previous), n);
  const auto r = Lotus::ReprojectTemporalHistory(*static_cast<const Lotus::TemporalPixel*>(current), x, y, width,
                                                 height, p, v);
  out[0] = r.valid;
  out[1] = static_cast<uint32_t>(r.rejection);
  out[2] = r.previous_index;
  out[3] = r.reserved;
}

// temporal.cpp:114-147 (TemporalHistory::Advance) over n frames' settings:
// each frame's TemporalResetReason, or 0xFFFFFFFF where Advance refused it.
void ref_temporal_advance(uint32_t n, const uint64_t* frame_index, const uint32_t* camera_cut,
                          const uint32_t* extent, const float* display, const int32_t* data,
                          const uint64_t* history_key, uint32_t* reasons) {
  Lotus::TemporalHistory history;
  const Lotus::FrameSnapshot snapshot{};
  for (uint32_t i = 0; i < n; ++i) {
    Lotus::TemporalFrameSettings s;
    s.frame_index = frame_index[i];
    s.camera_cut = camera_cut[i] != 0;
    s.width = extent[2 * i];
    s.height = extent[2 * i + 1];
    for (int k = 0; k < 4; ++k) {
      s.display_window[k] = display[4 * i + k];
      s.data_window[k] = data[4 * i + k];
    }
    s.history_key = history_key[i];
    try {
      reasons[i] = static_cast<uint32_t>(history.Advance(snapshot, s).reset_reason);
    } catch (const std::exception&) {
      reasons[i] = 0xFFFFFFFFu;
    }
  }
}

// temporal.cpp:514-528 (SettleTemporalRadiance) over `count` pixels: prior
// and still colours (4 floats each), each pixel's prior sample count.
void ref_settle(const float* prior, const float* still, const uint32_t* prior_samples, size_t count,
                uint32_t still_samples, uint32_t max_samples, float* out) {
  for (size_t i = 0; i < count; ++i) {
    std::array<float, 4> p, s;
    std::memcpy(p.data(), prior + 4 * i, 16);
    std::memcpy(s.data(), still + 4 * i, 16);
    const auto v = Lotus::SettleTemporalRadiance(p, s, prior_samples[i], still_samples, max_samples);
    std::memcpy(out + 4 * i, v.data(), 16);
  }
}

}  // extern "C"
