/* CPU oracle for the wavefront path tracer — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of hydra-lotus's reference transport (the Slang
 * reference integrator), used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker. Nothing in the product links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference repository root). Pinning: see the header comment of
 * lotus_oracle.cpp for which check pins which function.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The scene as the oracle reads it: object-space geometries placed by
 * instances, with one material record per slot. The material record is the
 * reference's GpuMaterialRecord (backend/vulkan/gpu_scene.hpp:94-113),
 * 36 words. */
/* An analytic light (Light, include/lotus/render_world.hpp:89-114): shape 0
 * sphere, 1 rect, 2 disk, 3 distant (angle 0: parallel). */
typedef struct oracle_light {
  uint32_t shape;
  uint32_t normalize;
  float world_from_light[16]; /* column-major */
  float radiance[3];
  float radius, width, height, angle;
} oracle_light;

/* A latitude-longitude environment image (EnvironmentMap,
 * include/lotus/render_world.hpp:116-133), RGBA32F rows from the top. */
typedef struct oracle_environment_map {
  uint32_t width, height;
  const float* texels;
  float scale[3];
  float world_from_environment[16];
} oracle_environment_map;

/* A material texture (Texture, include/lotus/material.hpp:15-41): format 0
 * RGBA8 linear, 1 RGBA8 sRGB, 2 RGBA32F; rows from the top. */
typedef struct oracle_texture {
  uint32_t width, height, format;
  const void* texels;
} oracle_texture;

typedef struct oracle_scene_desc {
  uint32_t geometry_count;
  const float* positions;           /* xyz per vertex, all geometries */
  const uint32_t* triangles;        /* 3 indices per triangle, geometry-relative */
  const uint32_t* geometry_vertex_offset;   /* per geometry */
  const uint32_t* geometry_triangle_offset; /* per geometry */
  const uint32_t* geometry_triangle_count;  /* per geometry */
  uint32_t instance_count;
  const uint32_t* instance_geometry; /* per instance */
  const float* instance_transform;   /* 16 floats per instance, column-major */
  const uint32_t* instance_material; /* per instance */
  uint32_t material_count;
  const uint32_t* materials;         /* 36 words per material */
  float environment[3];
  const uint32_t* geometry_vertex_count;    /* per geometry */
  uint32_t light_count;
  const oracle_light* lights;
  const oracle_environment_map* environment_map; /* NULL: none */
  uint32_t texture_count;
  const oracle_texture* textures;
  uint32_t binding_count;
  const uint32_t* bindings;                 /* 24 words per GpuTextureInputRecord */
  const float* texcoords;                   /* 2 per triangle corner */
  const uint32_t* geometry_texcoord_offset; /* per geometry: first corner or 0xFFFFFFFF; NULL: none */
  const uint32_t* instance_texcoord_offset; /* per instance (overrides the geometry's); NULL: none */
  const float* normals;                     /* authored shading normals, 3 per triangle corner */
  const uint32_t* geometry_normal_offset;   /* per geometry: first corner or 0xFFFFFFFF; NULL: none */
  const float* colors;                      /* vertex colours, 3 per entry */
  const uint32_t* instance_color_offset;    /* per instance: first entry or 0xFFFFFFFF; NULL: none */
  const uint32_t* instance_color_constant;  /* per instance: 1 one colour for the mesh, 0 per corner */
} oracle_scene_desc;

typedef struct oracle_scene oracle_scene;

oracle_scene* oracle_scene_create(const oracle_scene_desc* desc, char* error,
                                  size_t error_size);
void oracle_scene_destroy(oracle_scene* scene);
uint32_t oracle_scene_triangle_count(const oracle_scene* scene);
/* world-space corners of flattened triangle g, 9 floats */
void oracle_scene_world_triangles(const oracle_scene* scene, float* out);

/* Light table (backend/vulkan/gpu_scene.cpp:512-609): per instance first
 * entry or 0xFFFFFFFF, and per entry {pdf, threshold, alias, instance}. */
uint32_t oracle_light_count(const oracle_scene* scene);
void oracle_light_table(const oracle_scene* scene, uint32_t* offsets,
                        uint32_t* records /* 4 words per entry */);

/* Rays: origin xyz, tmin, direction xyz, tmax (8 floats). Hits: instance,
 * primitive (0xFFFFFFFF on a miss), barycentric u, v, t (5 words). */
void oracle_trace_closest(const oracle_scene* scene, uint32_t count,
                          const float* rays, uint32_t* hits, int brute_force,
                          int threads);
void oracle_trace_occluded(const oracle_scene* scene, uint32_t count,
                           const float* rays, uint32_t* occluded,
                           int brute_force, int threads);
/* ShadowTransmittance per ray (scene.slang:445-460). */
void oracle_trace_transmittance(const oracle_scene* scene, uint32_t count,
                                const float* rays, float* transmittance,
                                int brute_force, int threads);

typedef struct oracle_frame {
  float clip_to_world[16];  /* column-major */
  float ndc_from_pixel[4];
  uint32_t width, height;   /* target size; pixel index = y * width + x */
  uint32_t rect[4];         /* data window x, y, w, h */
  uint32_t sample_index;    /* first sample's random sequence */
  uint32_t spp;             /* samples per pixel this call */
  uint32_t max_bounces;
  uint32_t light_sampling;  /* 0 BSDF only, 1 NEE + MIS, 2 ReSTIR initial at the first hit */
  uint32_t restir_candidates; /* light draws per first hit (light_sampling 2) */
  uint32_t restir_proposal;   /* 0 power table, 1 uniform, 2 mixture */
} oracle_frame;

/* Renders spp samples per pixel of the data window. accum: 8 doubles per
 * target pixel: sum r, g, b, hit samples, sum r^2, g^2, b^2, 0. counts:
 * spp x (max_bounces + 2) ray-queue occupancies per sample (camera rays at
 * [0]); shadow: spp shadow-ray counts. */
int oracle_render(const oracle_scene* scene, const oracle_frame* frame,
                  double* accum, uint32_t* counts, uint64_t* shadow,
                  int threads);

/* One path's radiance, for single-path probes: rgb out, returns depth. */
uint32_t oracle_trace_path(const oracle_scene* scene, const oracle_frame* frame,
                           uint32_t x, uint32_t y, uint32_t sample_index,
                           float* rgb);

/* Reservoir arithmetic in the GPU's float form (reservoir.slang:23-103) on
 * 4-word reservoirs {sample, weight_sum, selected_target, M}: stream and
 * merge return 1 selected, 0 not, -1 rejected (reservoir unchanged). */
int oracle_reservoir_stream(uint32_t* reservoir, uint32_t sample, float target, float proposal, float random);
int oracle_reservoir_merge(uint32_t* destination, const uint32_t* source, float normalization, float target,
                           uint32_t cap, float random);
float oracle_reservoir_normalization(const uint32_t* reservoir);
float oracle_reuse_normalization(const uint32_t* reservoir, uint32_t supporting);

/* The initial reservoir of pixel (x, y) for a sample (restir.slang:163-233):
 * RestirPixel's 12 words {sample, weight_sum, target, M, u, v, proposal,
 * flags, contribution rgb, visibility}. Returns whether the camera ray hit. */
int oracle_restir_initial(const oracle_scene* scene, const oracle_frame* frame, uint32_t x, uint32_t y,
                          uint32_t sample_index, uint32_t* record);

/* ReSTIR DI reuse (restir.slang:302-491). A reuse record is RestirReusePixel's
 * 16 words {RestirPixel's 12, normalization, supporting candidates, previous
 * index, age}; a receiver surface is 40 floats: {offset origin, view depth},
 * {geometric normal, instance id low bits}, {shading normal, instance id high
 * bits}, {wo in the shading frame, c0}, {c1..c4}, {c5..c8}, {c9..c12},
 * {c13, c14, material id low, high}, {c15..c18}, {c19, c20, c21, model},
 * the closure words c being UsdPreviewSurface's {dielectric f90; base
 * colour, alpha; dielectric f0, metallic; f0, f90; transmission (negated:
 * thin wall), eta (negated: smooth)} and zeros (model 0), or glTF PBR's 22
 * packed terms (model 1).
 *
 * Temporal pass at pixel (x, y): the oracle's own initial reservoir and
 * receiver surface of the frame's camera ray, merged with the previous frame's
 * record and surface at previous_index when history_valid (the reprojection
 * accepted the pixel) and the pixel-centre surface's ids equal the receiver's
 * (TemporalSampleHistory, temporal.slang:74-79). view: world -> view,
 * column-major; ids NULL: instance + 1 and material slot + 1. Writes the
 * record and the receiver surface; returns whether the camera ray hit. */
int oracle_restir_temporal(const oracle_scene* scene, const oracle_frame* frame, uint32_t x, uint32_t y,
                           uint32_t sample_index, const float* view, const uint64_t* instance_ids,
                           const uint64_t* material_ids, uint32_t history_valid, uint32_t previous_index,
                           uint64_t centre_instance_id, uint64_t centre_material_id,
                           const uint32_t* previous_record, const float* previous_surface, uint32_t max_history,
                           uint32_t* record, float* surface);
/* Spatial pass at pixel (x, y) over the temporal pass's records and surfaces
 * of the whole width x height target: writes the record and the accepted
 * neighbours (5 words, 0xFFFFFFFF: none): the four axis neighbours at the
 * radius, or with `disk` five drawn within it. */
void oracle_restir_spatial(const oracle_scene* scene, uint32_t width, uint32_t height, uint32_t x, uint32_t y,
                           uint32_t sample_index, uint32_t radius, uint32_t max_history, uint32_t disk,
                           const uint32_t* records, const float* surfaces, uint32_t* record, uint32_t* sources);

/* The camera ray through a film position (reference CameraRay). */
void oracle_camera_ray(const float* clip_to_world, float ndc_x, float ndc_y,
                       float* ray /* 8 floats */);

/* OpenGL world_to_clip -> Vulkan world_to_clip -> inverse (float). Returns 0
 * on a singular matrix. */
int oracle_camera_constants(const float* world_to_clip_gl,
                            float* world_to_clip_vk, float* clip_to_world);

/* RNG probes (backend/vulkan/shaders/sampling/sampling.slang:9-25). */
/* EvaluatePreview's inputs at a hit: base colour, roughness, emission,
 * metallic, specular colour, 0 (12 floats), texture lookups applied. */
/* Render-buffer writes: a top-down product into a bottom-up buffer of the
 * same size (formats 0 UNorm8Vec4, 1 Float32Vec4, 2 Float32, 3 Int32);
 * 0 where the reference's WriteColor returns false. */
int oracle_write_render_buffer(const void* source, uint32_t source_format, uint32_t width, uint32_t height,
                               void* buffer, uint32_t buffer_format);
void oracle_material_inputs(const oracle_scene* scene, uint32_t instance, uint32_t primitive, float u, float v,
                            float* out);
uint32_t oracle_pcg_hash(uint32_t value);
uint32_t oracle_seed_random(uint32_t pixel, uint32_t sample);
void oracle_random_sequence(uint32_t state, uint32_t count, float* out);

/* Sine and cosine of 2 pi u, as every sampled angle uses them. */
void oracle_sincos_turns(uint32_t count, const float* u, float* s, float* c);

/* BSDF probes over one material record and a shading-frame direction pair.
 * eval: f*|cos| rgb; pdf; sample: wi xyz, weight rgb, pdf, delta, transmitted,
 * ok (10 floats). `entering` as EvaluatePreview decides it. */
void oracle_bsdf_eval(const uint32_t* material, int entering, const float* wo,
                      const float* wi, float* f, float* pdf);
void oracle_bsdf_sample(const uint32_t* material, int entering,
                        const float* wo, const float* u, float* out);

/* OffsetRayOrigin probe. */
void oracle_offset_ray_origin(const float* position, const float* normal,
                              float* out);

#ifdef __cplusplus
}
#endif
