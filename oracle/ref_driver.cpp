// C driver over the reference's own std-only CPU functions, compiled in place
// from the reference sources by `make -C oracle ref` into oracle/_ref/ (never
// committed). TEST INFRASTRUCTURE ONLY: tests use it to pin the oracle's
// restatements (fixtures, DES-Q5, reservoir arithmetic) against the
// reference's code where that code builds without Vulkan.
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include <lotus/render_world.hpp>
#include <lotus/temporal.hpp>

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

}  // extern "C"
