/* lotus_hip.h — C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * The drop-in boundary for hydra-lotus's Integrator::Wavefront path. Every
 * entry point replaces a reference interface, cited per function:
 *   include/lotus/vulkan_backend.hpp:753-883  OffscreenRenderer and its factory
 *   backend/vulkan/gpu_scene.hpp:248-510       GpuScene (reached via UpdateScene)
 *   backend/vulkan/shaders/common/scene.slang  Trace / ShadowTransmittance
 * Plain pointers and sizes only; no torch or HIP types cross it. Buffers are
 * caller-allocated. A handle is single-threaded (one frame in flight, as
 * vulkan_backend.hpp:744-752 states for the reference). Every call returns a
 * status: LOTUS_HIP_PASS, LOTUS_HIP_FAIL or LOTUS_HIP_SKIP (no device), with a
 * NUL-terminated detail string, mirroring FrameStatus {Pass, Fail, Skip}
 * (vulkan_backend.hpp:23-27). After a failed GPU submission the handle stays
 * failed ("the renderer failed earlier: ..."), as vulkan_backend.cpp:1095-1096
 * latches; invalid settings fail without latching.
 */
#ifndef LOTUS_HIP_H
#define LOTUS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LOTUS_HIP_PASS 0
#define LOTUS_HIP_FAIL 1
#define LOTUS_HIP_SKIP 77

/* The wavefront integrator traces at most this many bounces
 * (kMaxWavefrontBounces, vulkan_backend.hpp:694). */
#define LOTUS_HIP_MAX_BOUNCES 1024u

/* Hit word for a ray that found nothing. */
#define LOTUS_HIP_MISS 0xFFFFFFFFu

typedef struct lotus_hip_renderer lotus_hip_renderer;

/* The GPU scene's content, one plan replacing the previous one in full
 * (SceneUpdate, include/lotus/extraction.hpp:91-132, applied by
 * GpuScene::Apply, backend/vulkan/gpu_scene.cpp:1606): object-space
 * geometries placed by instances, and the material table as
 * GpuMaterialRecord words (gpu_scene.hpp:94-113, 36 words each). The
 * renderer compares a plan with the resident one to apply only what changed
 * (see lotus_hip_scene_evidence.update). */
/* An analytic light (Light, include/lotus/render_world.hpp:89-114). A
 * sphere is centred on its frame's origin, its radius scaled by the cube
 * root of the frame's absolute determinant; a rect or disk lies in its
 * frame's XY plane and emits from its -Z side; a distant light shines along
 * its frame's -Z axis from beyond the scene, `angle` its angular diameter in
 * degrees (0: a parallel light, whose radiance is its irradiance). */
#define LOTUS_HIP_LIGHT_SPHERE 0u
#define LOTUS_HIP_LIGHT_RECT 1u
#define LOTUS_HIP_LIGHT_DISK 2u
#define LOTUS_HIP_LIGHT_DISTANT 3u
typedef struct lotus_hip_light {
  uint32_t shape;             /* LOTUS_HIP_LIGHT_* */
  uint32_t normalize;         /* UsdLux normalize: radiance over area (or the distant light's size factor) */
  float world_from_light[16]; /* column-major */
  float radiance[3];          /* linear RGB, non-negative */
  float radius;               /* sphere, disk */
  float width, height;        /* rect, along X and Y */
  float angle;                /* distant, degrees */
} lotus_hip_light;

/* Texel formats of textures and environment images (TextureFormat,
 * include/lotus/material.hpp:15-41). */
#define LOTUS_HIP_TEXTURE_RGBA8_UNORM 0u
#define LOTUS_HIP_TEXTURE_RGBA8_SRGB 1u
#define LOTUS_HIP_TEXTURE_RGBA32F 2u

/* A latitude-longitude environment image (EnvironmentMap,
 * include/lotus/render_world.hpp:116-133, its texture any Texture format):
 * texels in `format` (LOTUS_HIP_TEXTURE_*, above; 8-bit sRGB channels
 * decode to linear values as TexelRgb does, gpu_scene.cpp:612-629), rows
 * from the top (+Y) to the bottom (-Y), column 0 at longitude +pi; its
 * radiance is the texels' RGB times `scale`, and only the frame's
 * orientation counts (its axes orthonormalized in X, Y, Z order). */
typedef struct lotus_hip_environment_map {
  uint32_t width, height;
  const void* texels;           /* 4 floats or 4 bytes per texel */
  float scale[3];
  float world_from_environment[16]; /* column-major */
  uint32_t format;              /* LOTUS_HIP_TEXTURE_RGBA32F, _RGBA8_SRGB or _RGBA8_UNORM */
} lotus_hip_environment_map;

/* A texture of the material bindings (Texture, include/lotus/material.hpp:
 * 15-41): texels row-major from the image's top; 8-bit sRGB channels are
 * decoded to linear values (alpha stays linear). */
typedef struct lotus_hip_texture {
  uint32_t width, height, format;
  const void* texels; /* 4 bytes or 4 floats per texel */
} lotus_hip_texture;

typedef struct lotus_hip_scene_plan {
  uint32_t geometry_count;
  const float* positions;                   /* xyz per vertex, all geometries */
  const uint32_t* triangles;                /* 3 geometry-relative indices per triangle */
  const uint32_t* geometry_vertex_offset;   /* per geometry, into positions */
  const uint32_t* geometry_vertex_count;    /* per geometry */
  const uint32_t* geometry_triangle_offset; /* per geometry, into triangles */
  const uint32_t* geometry_triangle_count;  /* per geometry */
  uint32_t instance_count;
  const uint32_t* instance_geometry;        /* per instance */
  const float* instance_transform;          /* 16 floats per instance, column-major */
  const uint32_t* instance_material;        /* per instance: material table slot */
  uint32_t material_count;
  const uint32_t* materials;                /* 36 words per material: UsdPreviewSurface
                                               (model 0) or glTF PBR (model 1), as
                                               FlattenMaterial writes them */
  float environment[3];                     /* constant environment radiance */
  uint32_t bvh_leaf_size;                   /* 0: default (2); at most 8 */
  uint32_t light_count;                     /* analytic lights, in the light table after the triangles */
  const lotus_hip_light* lights;
  const lotus_hip_environment_map* environment_map; /* NULL: none; added to the constant environment */
  /* Texture lookups of the materials (gpu_scene.hpp:56-84): the textures,
   * the binding table of GpuTextureInputRecord words (24 per binding; a
   * material record's words 1-3 give its first binding and its textured
   * slots), and per-corner texture coordinates (2 floats per triangle
   * corner) with each geometry's first corner, 0xFFFFFFFF for a geometry
   * without them. Under UsdPreviewSurface opacity, normal (tangent space),
   * base colour, roughness, metallic, emission and specular colour take
   * lookups; under glTF PBR its fourteen slots do. */
  uint32_t texture_count;
  const lotus_hip_texture* textures;
  uint32_t binding_count;
  const uint32_t* bindings;
  const float* texcoords;
  const uint32_t* geometry_texcoord_offset; /* per geometry; NULL: no geometry has texture coordinates */
  /* Per instance the first corner of the texture-coordinate set its material
   * reads (SceneInstance::texcoords, include/lotus/extraction.hpp:27-36),
   * 0xFFFFFFFF for none; NULL: each instance reads its geometry's
   * (geometry_texcoord_offset). */
  const uint32_t* instance_texcoord_offset;
  /* Authored shading normals (MeshGeometry::normals, render_world.hpp:43-55;
   * ShadingNormal, scene.slang:467-488): object-space normals, 3 floats per
   * triangle corner, and each geometry's first corner, 0xFFFFFFFF for a
   * geometry without them; NULL: none. */
  const float* normals;
  const uint32_t* geometry_normal_offset;
  /* Vertex colours (MeshGeometry::colors; HitColor, scene.slang:376-385),
   * 3 floats per entry, and per instance the first entry of the colour set
   * its material's base colour reads (SceneInstance::colors), 0xFFFFFFFF for
   * none (at most 2^31 - 1), with whether that set is one colour for the
   * whole mesh (1) or one per triangle corner (0); NULL: none. */
  const float* colors;
  const uint32_t* instance_color_offset;
  const uint32_t* instance_color_constant;
  /* How the BVH's topology is built (no reference counterpart: the
   * reference's driver builds its BLAS, gpu_scene.cpp:1899-1950):
   * LOTUS_HIP_BVH_LBVH (0), Morton order and radix tree on the GPU, the
   * fastest build; or LOTUS_HIP_BVH_SAH (1), binned SAH on the host's
   * threads, a slower build whose tree a ray walks through fewer nodes.
   * Boxes, the 4-wide collapse and refits run on the GPU either way. A
   * change of builder rebuilds. */
  uint32_t bvh_builder;
  /* What the caller vouches has not changed since the last plan it applied
   * (SceneUpdate's instances_changed, materials_changed, lights_changed,
   * environment_changed and its geometry and texture lists,
   * include/lotus/extraction.hpp:91-132): LOTUS_HIP_UNCHANGED_* bits. The
   * library then neither validates nor compares those parts again and keeps
   * what it holds for them; counts are always compared, and a renderer
   * without a scene ignores the bits. 0 checks everything. A plan whose
   * vouched parts did change renders the resident ones. */
  uint32_t unchanged;
  /* With LOTUS_HIP_UNCHANGED_POSITIONS: the geometries whose positions may
   * have changed (the others' are vouched for). */
  uint32_t moved_geometry_count;
  const uint32_t* moved_geometries;
} lotus_hip_scene_plan;

#define LOTUS_HIP_UNCHANGED_TOPOLOGY 1u   /* geometry offsets and counts, triangle indices, the instances'
                                             geometries, texture lookups (textures, bindings, corner sets) */
#define LOTUS_HIP_UNCHANGED_POSITIONS 2u  /* positions, but moved_geometries' */
#define LOTUS_HIP_UNCHANGED_TRANSFORMS 4u /* instance transforms */
#define LOTUS_HIP_UNCHANGED_MATERIALS 8u  /* the material table and the instances' slots */
#define LOTUS_HIP_UNCHANGED_LIGHTS 16u    /* analytic lights, the constant environment and the image */

#define LOTUS_HIP_BVH_LBVH 0u
#define LOTUS_HIP_BVH_SAH 1u

/* What a lotus_hip_scene_apply did, the counterpart of GpuScene::Apply's
 * retain / refit / rebuild decisions (gpu_scene.cpp:1638-1725; rules in
 * docs/reference/SCENE.md "Acceleration structures"). The plan is compared
 * with the resident one: */
#define LOTUS_HIP_SCENE_UNCHANGED 0u /* an equal plan: nothing is written, the
                                        accumulation continues (an empty
                                        SceneUpdate, vulkan_backend.cpp:2621-2626) */
#define LOTUS_HIP_SCENE_TABLES 1u    /* same geometry, indices, instances and
                                        transforms: the material table, the
                                        instances' material slots, the light
                                        table or the environment rewritten, the
                                        BVH kept */
#define LOTUS_HIP_SCENE_REFIT 2u     /* every geometry keeps its vertex count and
                                        exact triangle indices and every instance
                                        its geometry: changed positions and
                                        transforms re-place the triangles and
                                        refit the BVH's boxes in place (the
                                        BLAS point refit and TLAS update) */
#define LOTUS_HIP_SCENE_BUILD 3u     /* anything else (or the first plan): a new BVH */

/* GpuSceneEvidence / GpuSceneStats counterpart (vulkan_backend.hpp:412-506). */
typedef struct lotus_hip_scene_evidence {
  uint32_t triangle_count;
  uint32_t instance_count;
  uint32_t material_count;
  uint32_t light_count;      /* light table entries */
  uint32_t node_count;       /* wide BVH nodes */
  uint32_t bvh_depth;        /* wide levels, root = 1 */
  uint32_t leaf_count;
  uint32_t bvh_width;        /* children per node */
  double upload_ms;          /* host to device copies, wall */
  double build_gpu_ms;       /* this call's BVH build or refit kernels, hipEvent; 0 if neither ran */
  uint64_t device_bytes;     /* scene allocations */
  uint32_t update;           /* LOTUS_HIP_SCENE_* */
  uint32_t moved_geometries; /* geometries whose positions this call re-uploaded (refits) */
  /* Lifetime counts since the renderer was created (GpuSceneStats
   * blas_builds / blas_updates, material_writes, light_writes,
   * uploaded_bytes); an unchanged plan changes none of them. */
  uint64_t bvh_builds;
  uint64_t bvh_refits;
  uint64_t material_writes;
  uint64_t light_writes;
  uint64_t uploaded_bytes;
} lotus_hip_scene_evidence;

/* SceneOutput (vulkan_backend.hpp:624-637). */
#define LOTUS_HIP_OUTPUT_RADIANCE 0u
#define LOTUS_HIP_OUTPUT_BARYCENTRICS 1u
#define LOTUS_HIP_OUTPUT_SHADING_NORMAL 2u
#define LOTUS_HIP_OUTPUT_ALBEDO 3u
#define LOTUS_HIP_OUTPUT_ROUGHNESS_METALLIC 4u

/* LightSampling (vulkan_backend.hpp:652-674). */
#define LOTUS_HIP_LIGHT_BSDF 0u
#define LOTUS_HIP_LIGHT_MIS 1u
/* LightSampling::RestirInitial (vulkan_backend.hpp:650-660, scene.slang:249):
 * the first hit's direct light from a ReSTIR initial reservoir of
 * restir_candidates light draws (restir.slang:163-254); later hits as MIS. */
#define LOTUS_HIP_LIGHT_RESTIR 2u

/* One RenderScene call (vulkan_backend.hpp:789-791): DrawSummary::
 * world_to_clip, OffscreenTarget (vulkan_backend.hpp:253-275) and
 * PathTracingSettings (vulkan_backend.hpp:697-738). */
typedef struct lotus_hip_frame_params {
  float world_to_clip[16];    /* OpenGL clip space, column-major */
  uint32_t width, height;     /* target */
  /* OffscreenTarget's windows (vulkan_backend.hpp:253-262, resolved as
   * DisplayViewport and DataScissor, vulkan_backend.cpp:651-677): the
   * display window (x, y, w, h) maps pixels to NDC, the whole image when w
   * or h is not positive; the data window (x, y, w, h) is the pixels that
   * render, clipped to the image (outside it the clear colour stays), the
   * whole image when w or h is not positive. */
  float display_window[4];
  int32_t data_window[4];
  float clear_color[4];
  uint32_t frame_count;       /* samples this call adds */
  uint32_t output;            /* LOTUS_HIP_OUTPUT_* */
  uint32_t light_sampling;    /* LOTUS_HIP_LIGHT_* */
  uint32_t sample_index;      /* first sample's random sequence */
  uint32_t max_bounces;       /* scattering events after the camera hit */
  uint32_t max_samples;       /* 0 = no limit */
  uint32_t flags;             /* LOTUS_HIP_FRAME_* */
  uint32_t restir_candidates; /* LOTUS_HIP_LIGHT_RESTIR: light draws per first hit, 1..64 */
  uint32_t restir_proposal;   /* 0 power table, 1 uniform, 2 their mixture (scene.slang:919-943) */
} lotus_hip_frame_params;

/* Trace extension and shadow rays in separate launches, so each kind's
 * kernel time is its own (rays by kind); by default one launch per bounce
 * takes both. Results are the same either way. */
#define LOTUS_HIP_FRAME_SPLIT_KINDS 1u
/* Radiance cache frame (lotus_hip_cache_configure first): a subset of paths
 * trains the cache, every other path ends at the cache's query depth with a
 * cache query. */
#define LOTUS_HIP_FRAME_RADIANCE_CACHE 2u

/* GpuFrameEvidence subset (vulkan_backend.hpp:202-249) plus rays by kind.
 * Output buffers are optional (NULL skips the readback). */
typedef struct lotus_hip_frame_result {
  float* color;                 /* width*height*4, RGBA32F, rows from the top */
  float* depth;                 /* width*height, D32 (pixel-centre hit), 1 on a miss */
  uint32_t* path_counts;        /* capacity entries: last sample's ray-queue counts */
  uint32_t path_counts_capacity;
  uint32_t path_counts_size;    /* written: ends at the last bounce that traced */
  uint32_t wavefront_rounds;    /* bounces run as kernel rounds */
  uint32_t samples_per_pixel;   /* accumulated after the call */
  uint32_t frames_rendered;
  uint64_t camera_rays;         /* over all samples of the call */
  uint64_t extension_rays;
  uint64_t shadow_rays;
  double camera_gpu_ms;         /* WavefrontTimings (vulkan_backend.hpp:52-68) */
  double intersect_gpu_ms;      /* extension rays alone (split launches) */
  double shade_gpu_ms;
  double shadow_gpu_ms;         /* shadow rays alone (split launches) */
  double trace_gpu_ms;          /* extension + shadow rays in one launch */
  double tail_gpu_ms;           /* paths beyond the recorded rounds */
  double resolve_gpu_ms;
  double frame_gpu_ms;          /* first kernel start to last kernel end */
  /* Traversal work over the call's samples, per kernel: camera, extension
   * (intersect), shadow, tail. Wide nodes tested and triangles tested. */
  uint64_t traversal_nodes[4];
  uint64_t traversal_tris[4];
  /* The rays those traversals walked: camera, extension and shadow rays of
   * the rounds, and the tail's extension and shadow rays. */
  uint64_t traversal_rays[4];
  uint64_t trace_launches;      /* the rounds' traversal launches over the call's samples */
  /* Radiance cache frames: queries answered, training samples made, their
   * mean relative L2 loss before the frame's steps, and kernel times. */
  uint64_t cache_queries;
  uint64_t cache_samples;
  double cache_loss;
  double cache_infer_gpu_ms;
  double cache_train_gpu_ms;
  /* LOTUS_HIP_LIGHT_RESTIR: the last sample's reservoir records (RestirPixel,
   * vulkan_backend.hpp:152-164: 12 words per pixel {sample, weight_sum,
   * selected_target, M, u, v, proposal, flags, contribution rgb, visibility}),
   * optional, and the reservoir passes' GPU time. */
  uint32_t* restir_records;
  double restir_gpu_ms;
  /* Radiance cache frames: the last sample's training records (20 floats
   * each, the layout lotus_hip_cache_train takes), optional: at most
   * cache_training_capacity of them are written, their count in
   * cache_training_count. */
  float* cache_training;
  uint32_t cache_training_capacity;
  uint32_t cache_training_count;
  /* Temporal frames (TemporalProducts' pass timings,
   * vulkan_backend.hpp:120-146), hipEvent: the pixel-centre surface pass
   * with its reprojection, and the radiance accumulation, variance and
   * a-trous iterations (or a settling frame's crossfade); both are part of
   * frame_gpu_ms. */
  double temporal_surface_gpu_ms;
  double svgf_gpu_ms;
  /* Batched frames (lotus_hip_render_batch): 0 eager; with
   * LOTUS_HIP_BATCH_GRAPH, 1 when this frame's graph was instantiated, 2
   * when the last batch's executable graph of the same views took this
   * frame's parameters in place (hipGraphExecUpdate). */
  uint32_t batch_graph;
} lotus_hip_frame_result;

/* ProbeVulkanBackend (vulkan_backend.hpp:742): PASS when a gfx950 device is
 * visible, SKIP otherwise. */
int lotus_hip_probe(char* detail, size_t detail_size);

/* CreateOffscreenRenderer (vulkan_backend.hpp:879-883) on device `device`.
 * Returns NULL with *status SKIP (no device) or FAIL. */
lotus_hip_renderer* lotus_hip_create(int device, int* status, char* detail,
                                     size_t detail_size);
void lotus_hip_destroy(lotus_hip_renderer* renderer);

/* OffscreenRenderer::UpdateScene (vulkan_backend.hpp:816-817): uploads the
 * plan, builds the GPU scene (flattened world-space triangles, light table,
 * wide BVH built on the GPU) and returns after it is resident. */
int lotus_hip_scene_apply(lotus_hip_renderer* renderer,
                          const lotus_hip_scene_plan* plan,
                          lotus_hip_scene_evidence* evidence, char* detail,
                          size_t detail_size);

/* OffscreenRenderer::RenderScene with Integrator::Wavefront
 * (vulkan_backend.hpp:789-791). Accumulation restarts when the camera,
 * target, sample_index, max_bounces, light sampling or scene change
 * (vulkan_backend.cpp:1186-1205). */
int lotus_hip_render(lotus_hip_renderer* renderer,
                     const lotus_hip_frame_params* params,
                     lotus_hip_frame_result* result, char* detail,
                     size_t detail_size);

/* Scene.slang Trace (:424-439) and ShadowTransmittance (:445-460) over a
 * caller's rays: origin xyz, tmin, direction xyz, tmax (8 floats per ray).
 * Closest hits are instance, primitive (LOTUS_HIP_MISS on a miss),
 * barycentric weights of corners 1 and 2, t (5 words); any-hit results
 * (any_hit 1) are 1 (blocked: transmittance 0) or 0, and with any_hit 2
 * each ray's transmittance as a float's bits (the product of 1 - coverage
 * over the covered surfaces it crosses, 0 at an opaque one). Host
 * pointers. Ties in t resolve to the lowest flattened (instance,
 * primitive). A covered surface's candidate counts for a closest hit
 * with probability equal to its coverage, decided by a hash of the ray
 * and the candidate (the reference draws from the path's random stream in
 * the hardware's candidate order, which no other traversal reproduces). */
int lotus_hip_trace(lotus_hip_renderer* renderer, uint32_t count,
                    const float* rays, uint32_t* hits, int any_hit,
                    char* detail, size_t detail_size);

/* The same over device-resident rays and hits (hipMalloc'd by the caller),
 * enqueued on `stream` (a hipStream_t, NULL for the default stream);
 * returns without synchronizing. */
int lotus_hip_trace_device(lotus_hip_renderer* renderer, uint32_t count,
                           const void* rays_device, void* hits_device,
                           int any_hit, void* stream, char* detail,
                           size_t detail_size);

/* Launch shapes the autotuner chooses (no reference counterpart: the
 * reference's driver schedules its own dispatches). Results never depend on
 * them. */
typedef struct lotus_hip_tuning {
  uint32_t camera_dynamic; /* 1: a lane whose ray ends takes its next one without waiting for its wave; 0: the
                              wave's lanes take their next rays together */
  uint32_t trace_dynamic;  /* the same for extension and shadow rays */
  uint32_t tail_divisor;   /* bounces whose ray queue held at least 1/tail_divisor of the camera rays in the
                              previous sample run as wavefront rounds, the megakernel tail follows the rest
                              (vulkan_backend.cpp:369-385, whose divisor is 16, the default); 0: every bounce
                              that traced a ray runs as a round */
  uint32_t resident_percent; /* the persistent traversal kernels' (camera, trace, tail) workgroups as a
                                percentage of what their occupancy lets the device hold at once, in
                                [1, 100]: fewer workgroups each walk more rays; 0 means 100 */
  uint32_t megakernel;       /* 1: a frame of BSDF or MIS light sampling runs each path from its camera ray
                                to its end in one persistent launch (a lane pair per path, the tail's form)
                                instead of wavefront rounds; frames with the radiance cache, reservoirs,
                                temporal products, glTF PBR or several parallel lights under BSDF sampling
                                keep the rounds */
  uint32_t camera_tile;      /* the camera kernel's pixel tile per wave: 0 8 x 8, 1 16 x 4, 2 4 x 16 */
} lotus_hip_tuning;

int lotus_hip_set_tuning(lotus_hip_renderer* renderer, const lotus_hip_tuning* tuning, char* detail,
                         size_t detail_size);
int lotus_hip_get_tuning(lotus_hip_renderer* renderer, lotus_hip_tuning* tuning);

/* The light table as built (gpu_scene.cpp:512-609): per instance its first
 * entry or 0xFFFFFFFF, then 4 words per entry {pdf, threshold, alias,
 * instance}. Returns the entry count in *count; arrays may be NULL. */
int lotus_hip_light_table(lotus_hip_renderer* renderer, uint32_t* offsets,
                          uint32_t* records, uint32_t* count, char* detail,
                          size_t detail_size);

/* OffscreenRenderer::ReadBackScene (vulkan_backend.hpp:821, GpuSceneContents
 * :600-615): the resident scene as the device holds it (geometry, instance
 * records, material table, the wide BVH and its triangles), copied back.
 * Counts are always written; each array is filled when it is not NULL and
 * must hold what the counts say. Per-geometry vertex and triangle counts are
 * the applied plan's (the device keeps only offsets). */
typedef struct lotus_hip_scene_contents {
  uint32_t geometry_count;
  uint32_t instance_count;
  uint32_t material_count;
  uint32_t triangle_count;    /* flattened (instance, primitive) triangles */
  uint32_t node_count;        /* wide BVH nodes */
  uint32_t light_count;       /* light table entries */
  uint64_t position_floats;   /* length of `positions` (3 per vertex) */
  uint64_t index_words;       /* length of `triangles` (3 per triangle) */
  float environment[3];
  float* positions;
  uint32_t* triangles;
  uint32_t* geometry_vertex_offset;   /* geometry_count each */
  uint32_t* geometry_vertex_count;
  uint32_t* geometry_triangle_offset;
  uint32_t* geometry_triangle_count;
  uint32_t* instance_geometry;        /* instance_count */
  float* instance_transform;          /* 16 per instance, column-major */
  uint32_t* instance_material;        /* instance_count */
  uint32_t* materials;                /* 36 words per material */
  uint32_t* bvh_nodes;                /* 32 words per node: child boxes lo/hi x, y, z (4 each), children, padding */
  float* bvh_triangles;               /* 12 per triangle in leaf order: corners xyz with the flattened id,
                                         instance and primitive as the w words' bits */
} lotus_hip_scene_contents;

int lotus_hip_readback_scene(lotus_hip_renderer* renderer, lotus_hip_scene_contents* contents, char* detail,
                             size_t detail_size);

/* The accumulation as sums: 4 floats per pixel (rgb, hit samples), rows from
 * the top, and the count of samples added. */
int lotus_hip_read_accumulation(lotus_hip_renderer* renderer, float* sums,
                                uint32_t* samples, char* detail,
                                size_t detail_size);

/* Radiance cache: a 16-level hash-grid encoding (2 FP16 features per level)
 * and a 4 x 64 ReLU MLP without biases, trained online with Adam on the
 * relative L2 loss (pred - y)^2 / (pred^2 + 0.01) (Mueller et al. 2021,
 * 2022). No reference counterpart: the reference lists a neural radiance
 * cache only as a roadmap candidate (docs/design/ROADMAP_POLICY.md:526-535).
 *
 * Queries are 16 floats: position mapped to [0, 1]^3, unit direction towards
 * the previous path vertex, shading normal, roughness, base colour (3),
 * metallic, 2 unused. Training samples are a query, an RGB target and 1
 * unused float (20 floats). Parameters are FP32 masters: weights [16384] =
 * W0[64][48], W1..W3[64][64], W4[16][64] (rows 3..15 unused), the grid
 * [16][2^log2_table][2]. */
typedef struct lotus_hip_cache_config {
  uint32_t log2_table;      /* entries per level, 2^log2_table, 10..22 */
  uint32_t min_resolution;  /* cells per axis of the coarsest level */
  uint32_t max_resolution;  /* cells per axis of the finest level */
  uint32_t batch_capacity;  /* most samples of one optimizer step */
  float learning_rate, beta1, beta2, epsilon;
  uint32_t seed;            /* parameter initialisation */
  uint32_t query_depth;     /* frames: bounce at which other paths query (>= 1) */
  uint32_t train_paths;     /* frames: training paths per frame */
  uint32_t train_steps;     /* frames: optimizer steps per frame */
  uint32_t resolution[16];  /* written: cells per axis of each level */
  uint32_t dense_levels;    /* written: bit l set when level l is indexed densely */
} lotus_hip_cache_config;

/* (Re)creates the cache with freshly initialised parameters. */
int lotus_hip_cache_configure(lotus_hip_renderer* renderer, lotus_hip_cache_config* config, char* detail,
                              size_t detail_size);

/* radiance[i][3] = network(queries[i][16]) for i < count; device pointers,
 * synchronous. */
int lotus_hip_cache_infer(lotus_hip_renderer* renderer, const void* queries_device, uint32_t count,
                          void* radiance_device, char* detail, size_t detail_size);

/* Mean gradients of the loss over samples[count][20] (device pointer) without
 * a step: weights [16384] and grid [16][2^log2_table][2] (host arrays, either
 * may be NULL), and the mean loss. */
int lotus_hip_cache_gradients(lotus_hip_renderer* renderer, const void* samples_device, uint32_t count,
                              float* weight_grad, float* grid_grad, float* loss, char* detail,
                              size_t detail_size);

/* One Adam step over samples[count][20] (device pointer); `loss` (may be
 * NULL) receives the mean loss before the step. The grid's moments advance
 * only where a sample touched it. */
int lotus_hip_cache_train(lotus_hip_renderer* renderer, const void* samples_device, uint32_t count, float* loss,
                          char* detail, size_t detail_size);

/* Reads (write = 0) or writes (write = 1) the FP32 parameters through host
 * arrays (either may be NULL). Writing resets the Adam moments and step. */
int lotus_hip_cache_parameters(lotus_hip_renderer* renderer, float* weights, float* grid, int write, char* detail,
                               size_t detail_size);

/* Temporal frames: OffscreenRenderer::RenderTemporalScene
 * (vulkan_backend.hpp:807-810) with TemporalFrameSettings
 * (include/lotus/temporal.hpp:58-103). One call is one temporal frame: a
 * fresh batch of frame_count samples, the pixel-centre surface product and
 * its reprojection onto the previous frame (temporal.slang:497-627), ReSTIR
 * DI reuse (restir.slang:302-491) and SVGF (temporal.slang:229-493). */
typedef struct lotus_hip_temporal_settings {
  uint64_t frame_index;          /* must increase; a gap drops the history */
  uint64_t history_key;          /* changing it drops the history */
  uint32_t camera_cut;
  uint32_t accumulate_radiance;  /* SVGF temporal accumulation */
  uint32_t max_history_samples;  /* [1, 65536] */
  uint32_t filter_iterations;    /* a-trous iterations, at most 8 */
  float view[16];                /* world -> view of this frame's camera, column-major */
  float validation[4];           /* depth abs / rel tolerance, min normal cosine, max motion px; zeros = defaults */
  uint32_t restir_temporal;      /* PathTracingSettings::restir_temporal (vulkan_backend.hpp:719-724) */
  uint32_t restir_spatial;
  uint32_t restir_spatial_radius;          /* [1, 32] */
  uint32_t restir_max_history_candidates;  /* [1, 65536] */
  /* Optional per-instance identity (temporal.slang:32-40), host arrays of
   * the scene's instance_count entries, copied to the device each frame;
   * NULL: instance + 1, material slot + 1, rigid and unchanged with a static
   * placement. */
  const uint64_t* instance_ids;
  const uint64_t* material_ids;
  const uint32_t* instance_flags;          /* bit 1 rigid correspondence, bit 2 material unchanged */
  const float* previous_from_current;      /* 16 floats per instance, column-major */
  /* Optional products, host buffers of width*height records in the
   * reference's layouts (temporal.hpp:193-245): TemporalPixel (64 B),
   * TemporalValidityPixel (16 B), TemporalRadiancePixel (32 B). */
  void* surface;
  void* validity;
  void* radiance;
  uint32_t reset_reason;       /* written: TemporalResetReason (temporal.hpp:16-26) */
  uint32_t history_available;  /* written */
  /* Optional ReSTIR products (RestirProducts, vulkan_backend.hpp:175-200),
   * host buffers of width*height records: the temporal pass's and the
   * spatial pass's RestirReusePixel (64 B: the RestirPixel, normalization,
   * supporting candidates, previous index, age) and the spatial pass's
   * accepted neighbours (4 pixel indices, 0xFFFFFFFF for none; see
   * restir_spatial_neighbours for the disk mode). Written when
   * the frame ran that pass. */
  void* restir_temporal_records;
  void* restir_spatial_records;
  void* restir_spatial_sources;
  /* Optional: the reuse passes' receiver surfaces (RestirSurface,
   * restir.slang:30-62, in this renderer's 160-byte form of 10 float4:
   * {offset origin, view depth}, {geometric normal, instance id low bits},
   * {shading normal, instance id high bits}, {wo in the shading frame, c0},
   * {c1..c4}, {c5..c8}, {c9..c12}, {c13, c14, material id low, high bits},
   * {c15..c18}, {c19, c20, c21, model}; zero where the camera ray missed),
   * width*height of them. The closure words c of UsdPreviewSurface (model
   * 0): dielectric f90; base colour, alpha; dielectric f0, metallic; f0,
   * f90; transmission (negated: thin wall), eta (negated: smooth); zeros.
   * Of glTF PBR (model 1): its 22 terms in PackGltfPbrTerms' order
   * (gltf_pbr.slang:159-190). Written when the frame ran a reuse pass. */
  void* restir_surfaces;
  /* The spatial pass's neighbours: LOTUS_HIP_SPATIAL_AXES, the reference's
   * four at restir_spatial_radius along the axes (restir.slang:408-410), or
   * LOTUS_HIP_SPATIAL_DISK, five drawn uniformly within the radius from a
   * random stream of their own (no reference counterpart). In the disk mode
   * restir_spatial_sources holds 8 words per pixel, the five draws in order
   * and three 0xFFFFFFFF. */
  uint32_t restir_spatial_neighbours;
  /* TemporalHistoryRectification (include/lotus/temporal.hpp:35-57), which
   * needs accumulate_radiance: LOTUS_HIP_RECTIFY_NONE, _CLAMP (the
   * reprojected illumination clamped to the 3x3 fresh neighbourhood's mean
   * plus or minus one standard deviation), _ANTI_LAG (the kept history
   * shrunk by saturate((z - 2) / 2) of the 5x5 neighbourhood's luminance
   * difference in standard errors) or _REFLECTION (the kept history scaled
   * by 1 - s + s / (1 + (theta / a)^2) of the point's view-direction
   * change). Changing it resets no history. */
  uint32_t history_rectification;
  /* TemporalFrameSettings::settle (temporal.hpp:77-88), which needs
   * accumulate_radiance: once nothing changes (no reset, the same scene
   * generation, view and world_to_clip, no reservoir reuse, and the
   * preceding frame also asked to settle), frames stop accumulating history
   * and filtering and continue a progressive accumulation of still samples
   * from the host's sample_index (at most max_samples of them), whose mean
   * the output crossfades into from the last changing frame's colour over
   * max_history_samples (capped by max_samples) samples where that frame's
   * history had samples (SettleTemporalRadiance, temporal.slang:477-493). */
  uint32_t settle;
  /* Written (TemporalProducts, vulkan_backend.hpp:120-146): whether the
   * frame settled and its still samples, the first sample index of the
   * samples it added and their count. Frames with a fresh batch (radiance
   * accumulation or reservoir reuse) that do not settle draw their batch
   * from sample_index plus the samples of the preceding frames since the
   * history was last dropped (vulkan_backend.cpp:1805-1813). */
  uint32_t settling;
  uint32_t settled_samples;
  uint32_t batch_sample_index;
  uint32_t batch_samples;
  /* Optional host buffer of width*height float4: the frame's resolved
   * colour before accumulation (the fresh batch), or on a settling frame
   * the still samples' mean. */
  float* batch;
} lotus_hip_temporal_settings;

#define LOTUS_HIP_RECTIFY_NONE 0u
#define LOTUS_HIP_RECTIFY_CLAMP 1u
#define LOTUS_HIP_RECTIFY_ANTI_LAG 2u
#define LOTUS_HIP_RECTIFY_REFLECTION 3u

#define LOTUS_HIP_SPATIAL_AXES 0u
#define LOTUS_HIP_SPATIAL_DISK 1u

int lotus_hip_render_temporal(lotus_hip_renderer* renderer, const lotus_hip_frame_params* params,
                              lotus_hip_temporal_settings* temporal, lotus_hip_frame_result* result, char* detail,
                              size_t detail_size);

/* SVGF's variance estimate and filter over caller device buffers of
 * width*height TemporalPixel and TemporalRadiancePixel records: writes each
 * radiance record's variance word (EstimateTemporalVariance,
 * temporal.cpp:375-421) and, for iterations > 0, the filtered, remodulated
 * colour (FilterTemporalRadiance, temporal.cpp:424-512) as RGBA32F. */
int lotus_hip_svgf(lotus_hip_renderer* renderer, const void* surface_device, void* radiance_device, uint32_t width,
                   uint32_t height, uint32_t batch_samples, uint32_t max_history_samples, uint32_t iterations,
                   void* color_device, char* detail, size_t detail_size);

/* AccumulateTemporalRadiance (temporal.slang:219-267) over caller device
 * buffers of width*height records: each pixel's fresh batch colour (RGBA32F),
 * TemporalPixel, TemporalValidityPixel and the previous frame's
 * TemporalRadiancePixel records into this frame's TemporalRadiancePixel,
 * with fresh_samples samples in the batch, at most max_history_samples kept,
 * under a LOTUS_HIP_RECTIFY_* mode (the reflection weight reads each
 * surface's reserved word). */
int lotus_hip_accumulate(lotus_hip_renderer* renderer, const void* batch_device, const void* surface_device,
                         const void* validity_device, const void* old_radiance_device, uint32_t width,
                         uint32_t height, uint32_t fresh_samples, uint32_t max_history_samples,
                         uint32_t rectification, void* radiance_device, char* detail, size_t detail_size);

/* lotus_hip_svgf without allocating: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same kernels and
 * results as lotus_hip_svgf, with the filter's scratch (64 B per pixel) in a
 * buffer of the library's own, compiled into it, of at most 2,097,152 pixels
 * (1920 x 1080 fits); a larger image fails with a message and nothing
 * written. Calls no hipMalloc, hipFree or allocating routine. Every call on
 * one GPU shares that GPU's buffer, so calls that overlap on one GPU (from
 * several renderers or threads) race; the buffer keeps the last call's
 * contents, which no call reads before writing. */
int lotus_hip_svgf_static(lotus_hip_renderer* renderer, const void* surface_device, void* radiance_device,
                          uint32_t width, uint32_t height, uint32_t batch_samples, uint32_t max_history_samples,
                          uint32_t iterations, void* color_device, char* detail, size_t detail_size);

/* History reprojection over caller device buffers (ReprojectTemporalHistory,
 * temporal.cpp:258-300): TemporalValidityPixel records from current and
 * previous TemporalPixel records (previous NULL: no history); validation
 * holds the four tolerances of lotus_hip_temporal_settings, NULL for the
 * defaults. */
int lotus_hip_reproject(lotus_hip_renderer* renderer, const void* current_device, const void* previous_device,
                        uint32_t width, uint32_t height, const float* validation, void* validity_device,
                        char* detail, size_t detail_size);

/* Batched frames (no reference counterpart: the reference renders one target
 * per call with a blocking submit, vulkan_backend.cpp:1238-1276). View i
 * renders one radiance sample through renderers[i] (each view its own
 * renderer and scene; BSDF or MIS light sampling) under the accumulation
 * rules of lotus_hip_render, every view on its own stream with no host
 * synchronisation between views. LOTUS_HIP_BATCH_GRAPH captures the whole
 * frame, one branch per view, as one hipGraph and launches it. results
 * (optional) receive each view's counts and colour; gpu_ms the frame's GPU
 * time. Results are the same as unbatched frames. */
#define LOTUS_HIP_BATCH_GRAPH 1u
#define LOTUS_HIP_BATCH_FUSED 2u /* one path megakernel for every view, fed by one device work queue of
                                    256-path chunks (views of the same kernel variant share a launch);
                                    every view's frame must be one the megakernel runs; eager only */
int lotus_hip_render_batch(lotus_hip_renderer* const* renderers, const lotus_hip_frame_params* params,
                           lotus_hip_frame_result* results, uint32_t count, uint32_t flags, double* gpu_ms,
                           char* detail, size_t detail_size);

/* Autotuner (no reference counterpart): times trials frames of `params`
 * under each combination of lotus_hip_tuning's choices (the traversal
 * drivers, the tail rule, the path megakernel, 100 or 50 percent resident
 * workgroups), then the camera tile shapes for the chosen one, and keeps the
 * fastest; writes it to `chosen`. Trial frames do not join an
 * accumulation, and the next frame restarts it. */
int lotus_hip_autotune(lotus_hip_renderer* renderer, const lotus_hip_frame_params* params, uint32_t trials,
                       lotus_hip_tuning* chosen, char* detail, size_t detail_size);

/* Render-buffer writes of the products, on the device. Formats of a product
 * and of a host's render buffer, as HdFormat names them. The writes run on
 * the renderer's own stream, which does not wait for other streams: work that
 * fills a device buffer a call reads must have completed before the call. */
#define LOTUS_HIP_BUFFER_UNORM8X4 0u  /* HdFormatUNorm8Vec4 ("rgba8-unorm") */
#define LOTUS_HIP_BUFFER_FLOAT32X4 1u /* HdFormatFloat32Vec4 ("rgba32-sfloat") */
#define LOTUS_HIP_BUFFER_FLOAT32 2u   /* HdFormatFloat32 (depth) */
#define LOTUS_HIP_BUFFER_INT32 3u     /* HdFormatInt32 (ids) */

/* Replaces HdLotusRenderBuffer::WriteColor and WriteDepth
 * (adapters/hydra2/src/adapter.cpp:2544-2596) with WriteRowsFlippedLocked
 * (:2713-2731): copies a top-down product of source_width x source_height
 * pixels into a bottom-up render buffer on the device. A colour product
 * converts between RGBA32F and UNorm8 (lround(clamp(v, 0, 1) * 255), and
 * v / 255); every other pair must be one format. A buffer of another size, or
 * a pair it does not convert, fails without writing, as WriteColor returns
 * false. Ordered on the renderer's stream; returns once written. */
int lotus_hip_write_render_buffer(lotus_hip_renderer* renderer, const void* product_device, uint32_t product_format,
                                  uint32_t source_width, uint32_t source_height, void* buffer_device,
                                  uint32_t buffer_format, uint32_t buffer_width, uint32_t buffer_height,
                                  char* detail, size_t detail_size);

/* Replaces HdLotusRenderBuffer::WriteIds (adapter.cpp:2733-2743): fills an
 * Int32 render buffer of width x height pixels with `value`. */
int lotus_hip_write_ids(lotus_hip_renderer* renderer, void* buffer_device, uint32_t buffer_format, uint32_t width,
                        uint32_t height, int32_t value, char* detail, size_t detail_size);

/* The named temporal AOVs (adapter.cpp:275-285): lotus:motion (Float32Vec4,
 * the surface record's motion and depth), lotus:historyLength (Int32, the
 * radiance record's sample count, 0 without radiance records),
 * lotus:historyValidity and lotus:historyRejection (Int32, the validity
 * record's valid and rejection words). */
#define LOTUS_HIP_AOV_MOTION 0u
#define LOTUS_HIP_AOV_HISTORY_LENGTH 1u
#define LOTUS_HIP_AOV_HISTORY_VALIDITY 2u
#define LOTUS_HIP_AOV_HISTORY_REJECTION 3u

/* Replaces HdLotusRenderBuffer::WriteTemporal (adapter.cpp:2598-2630) on the
 * device: from source_width x source_height top-down device records,
 * TemporalPixel (64 B), TemporalValidityPixel (16 B) and
 * TemporalRadiancePixel (32 B, may be NULL), writes the AOV into a bottom-up
 * render buffer of its format (Float32Vec4 for motion, else Int32). A buffer
 * of another format or size fails without writing. Returns once written. */
int lotus_hip_write_temporal_aov(lotus_hip_renderer* renderer, uint32_t aov, const void* surface_device,
                                 const void* validity_device, const void* radiance_device, uint32_t source_width,
                                 uint32_t source_height, void* buffer_device, uint32_t buffer_format,
                                 uint32_t buffer_width, uint32_t buffer_height, char* detail, size_t detail_size);

/* The named ReSTIR AOVs (adapter.cpp:256-268): lotus:restirLight,
 * lotus:restirCandidates, lotus:restirSupport, lotus:restirAge (Int32),
 * lotus:restirWeight and lotus:restirDirect (Float32Vec4). */
#define LOTUS_HIP_AOV_RESTIR_LIGHT 0u
#define LOTUS_HIP_AOV_RESTIR_CANDIDATES 1u
#define LOTUS_HIP_AOV_RESTIR_SUPPORT 2u
#define LOTUS_HIP_AOV_RESTIR_AGE 3u
#define LOTUS_HIP_AOV_RESTIR_WEIGHT 4u
#define LOTUS_HIP_AOV_RESTIR_DIRECT 5u

/* Replaces HdLotusRenderBuffer::WriteRestir (adapter.cpp:2632-2672) on the
 * device: per pixel the resolved RestirReusePixel (64 B; the spatial pass's
 * records, else the temporal pass's; may be NULL), else the initial
 * RestirPixel (48 B; may be NULL), else an empty record (light 0xFFFFFFFF,
 * zeros). lotus:restirWeight writes {weight_sum, selected_target,
 * normalization, proposal}, the normalization the reuse record's, else
 * (weight_sum / M) / selected_target when M > 0 and the target is positive,
 * else 0; lotus:restirSupport the reuse record's supporting candidates, else
 * M; lotus:restirAge the reuse record's age, else 0. A buffer of another
 * format or size fails without writing. Returns once written. */
int lotus_hip_write_restir_aov(lotus_hip_renderer* renderer, uint32_t aov, const void* initial_device,
                               const void* resolved_device, uint32_t source_width, uint32_t source_height,
                               void* buffer_device, uint32_t buffer_format, uint32_t buffer_width,
                               uint32_t buffer_height, char* detail, size_t detail_size);

#ifdef __cplusplus
}
#endif

#endif /* LOTUS_HIP_H */
