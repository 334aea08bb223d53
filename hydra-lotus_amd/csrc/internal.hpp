// Internal host/device interfaces of liblotus_hip: the resident GPU scene and
// the launchers each translation unit exports to the host orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "lotus_hip.h"
#include "scene.hpp"

namespace lotus {

// One emissive triangle of the light table (GpuLightRecord, gpu_scene.hpp:120-126).
struct LightRecord {
  float pdf;
  float threshold;
  uint32_t alias;
  uint32_t instance;
};

// A light-table entry whose `instance` has this bit is an analytic light,
// the other bits its record's index (gpu_scene.hpp:128-131).
constexpr uint32_t kAnalyticLight = 0x80000000u;

// AnalyticLightRecord shapes (gpu_scene.hpp:136-142).
constexpr uint32_t kSphereLight = 0, kRectLight = 1, kDiskLight = 2, kDistantLight = 3, kParallelLight = 4;

// One analytic light in world space (GpuAnalyticLightRecord,
// gpu_scene.hpp:144-166): position (sphere radius in w), half axes u and v
// (a distant light's unit direction towards it in u, its cone's
// 1 - cos(theta_max) in u.w), the emitting side's unit normal (area in w),
// radiance (the light table's probability of the light in w), shape.
struct alignas(16) AnalyticRecord {
  float position[4];
  float axis_u[4];
  float axis_v[4];
  float normal[4];
  float radiance[4];
  uint32_t shape;
  uint32_t reserved[3];
};
static_assert(sizeof(AnalyticRecord) == 96, "analytic light record is 96 bytes");

// One interior node of the binary crossing hierarchy over the sphere, rect
// and disk lights (GpuLightNode, gpu_scene.hpp:175-185): per child its
// bounds, the child word in low.w (kLightLeaf | light, a node, or
// kNoLightChild).
struct alignas(16) LightNode {
  float low0[4], high0[4], low1[4], high1[4];
};
static_assert(sizeof(LightNode) == 64, "light node is 64 bytes");
constexpr uint32_t kLightLeaf = 0x80000000u;
constexpr uint32_t kNoLightChild = 0xFFFFFFFFu;
constexpr int kLightStack = 32;  // a median split's depth is at most ceil(log2(count))

// Everything the kernels read of the scene, as device pointers.
struct SceneView {
  const WideNode* nodes;
  const Tri* tris;                  // BVH order
  const uint32_t* tri_of_global;    // flattened id -> BVH order
  const uint32_t* instance_base;    // first flattened id per instance
  const uint32_t* instance_material;
  const float* instance_transform;  // 16 per instance
  const uint32_t* materials;        // 36 words per material
  const uint32_t* light_offsets;    // per instance
  const LightRecord* lights;
  const AnalyticRecord* analytic;   // analytic_count records
  const LightNode* light_nodes;     // light_node_count nodes, then the distant lights' indices
  const uint32_t* distant;          // distant_count indices of distant (cone) lights
  uint32_t light_count;
  uint32_t tri_count;
  uint32_t instance_count;
  uint32_t analytic_count, light_node_count, distant_count;
  uint32_t parallel;                // the first parallel light's index + 1 (BSDF sampling traces it), 0 without one
  uint32_t parallel_count;          // parallel lights (more than one: BSDF sampling walks them inline)
  uint32_t gltf_materials;          // materials of the glTF PBR model
  const float* albedo_tables;       // the glTF closure's directional albedo tables (albedo_tables.inc)
  float environment[3];
  // The environment image (env_width 0: none): texels, then its sampling
  // distribution's floats (cell densities, the rows' cumulative
  // distribution, each row's), the frame's axes as rows, the scale, and the
  // probability its light-table entry has.
  const float4* env_texels;
  const float* env_dist;
  uint32_t env_width, env_height, env_columns, env_rows;
  float env_frame[9];
  float env_scale[3];
  float env_selection;
  // Material texture lookups: each texture's {first texel, width, height},
  // the texels as linear RGBA, the binding table (24 words per binding),
  // per-corner texture coordinates and each instance's first corner
  // (0xFFFFFFFF: none).
  const uint4* textures;
  const float4* texels;
  const uint32_t* bindings;
  const float2* texcoords;
  const uint32_t* instance_texcoords;
  // Authored shading normals (3 floats per corner) and vertex colours (3
  // floats per entry): per instance its first corner, 0xFFFFFFFF for none
  // (a colour set's bit 31: one colour for the whole mesh). Null arrays: no
  // instance has them.
  const float* normals;
  const uint32_t* instance_normals;
  const float* colors;
  const uint32_t* instance_colors;
  // Instances whose material's coverage is tested (their triangles carry
  // kCoveredBit); 0: every surface is opaque.
  uint32_t covered_instances;
};

// The light-table entry of the environment image (gpu_scene.hpp:132-134).
constexpr uint32_t kEnvironmentLight = 0xFFFFFFFFu;

// Whether the kernels' variants with analytic lights, the environment image,
// material texture lookups, texture coordinates, authored normals and vertex
// colours run (the others compile none of that code).
__host__ __device__ inline bool scene_has_extras(const SceneView& S) {
  return S.analytic_count != 0 || S.env_width != 0 || S.bindings != nullptr || S.texcoords != nullptr ||
         S.instance_normals != nullptr || S.instance_colors != nullptr || S.covered_instances != 0 ||
         S.gltf_materials != 0;
}

// Device memory of the resident scene, kept until a plan needs a new BVH.
struct SceneBuffers {
  WideNode* nodes = nullptr;
  Tri* tris = nullptr;
  uint32_t* tri_of_global = nullptr;
  uint32_t* instance_base = nullptr;
  uint32_t* instance_material = nullptr;
  float* instance_transform = nullptr;
  uint32_t* materials = nullptr;
  uint32_t* light_offsets = nullptr;
  LightRecord* lights = nullptr;
  AnalyticRecord* analytic = nullptr;
  LightNode* light_nodes = nullptr;
  uint32_t* distant = nullptr;
  uint32_t analytic_count = 0, light_node_count = 0, distant_count = 0, parallel_count = 0, parallel = 0;
  uint32_t gltf_materials = 0;  // materials of the glTF PBR model
  float4* env_texels = nullptr;
  float* env_dist = nullptr;
  uint32_t env_width = 0, env_height = 0, env_columns = 0, env_rows = 0;
  float env_frame[9] = {};
  float env_scale[3] = {};
  float env_selection = 0.0f;
  uint4* textures = nullptr;
  float4* texels = nullptr;
  uint32_t* bindings = nullptr;
  float2* texcoords = nullptr;
  uint32_t* instance_texcoords = nullptr;
  float* normals = nullptr;
  uint32_t* instance_normals = nullptr;
  float* colors = nullptr;
  uint32_t* instance_colors = nullptr;
  // Per instance 1 when its material's coverage is tested (the host's copy
  // beside it); covered_instances counts them.
  uint32_t* instance_covered = nullptr;
  std::vector<uint32_t> covered_host;
  uint32_t covered_instances = 0;
  uint64_t normal_count = 0, color_count = 0;
  uint32_t texture_count = 0, binding_count = 0;
  uint64_t texel_count = 0, texcoord_count = 0;
  // The plan's geometry as uploaded: refits re-place its triangles and
  // read-back copies it.
  float* positions = nullptr;
  uint32_t* triangles = nullptr;
  uint32_t* geometry_vertex_offset = nullptr;
  uint32_t* geometry_triangle_offset = nullptr;
  uint32_t* instance_geometry = nullptr;
  uint64_t vertex_words = 0, index_words = 0;
  uint32_t material_count = 0, light_count = 0;
  uint32_t node_count = 0;
  uint32_t depth = 0;
  uint32_t leaf_count = 0;
  // Wide nodes per level, the root's first: the build numbers nodes level by
  // level, so level d holds the level_sizes[d] nodes after the levels above.
  std::vector<uint32_t> level_sizes;
  // Refit scratch, allocated by the first refit: each node's box before the
  // margin (lo.xyz, hi.xyz as two float4).
  float4* exact = nullptr;
  uint64_t bytes = 0;
};

// Inputs of the GPU build, already on the device.
struct BuildInput {
  const float* positions;
  const uint32_t* triangles;
  const uint32_t* geometry_vertex_offset;
  const uint32_t* geometry_triangle_offset;
  const uint32_t* instance_geometry;
  const float* instance_transform;
  const uint32_t* instance_base;  // prefix sum of instance triangle counts, instance_count + 1
  const uint32_t* instance_covered;  // per instance 1 when its coverage is tested; null: none is
  uint32_t instance_count;
  uint32_t tri_count;
  uint32_t leaf_size;
  uint32_t builder;  // LOTUS_HIP_BVH_LBVH or LOTUS_HIP_BVH_SAH
};

// The binary tree's topology by binned SAH on the host (sah_build.hip):
// leaf order, children and leaf ranges in bvh_build.hip's format.
bool sah_topology(const Tri* world, uint32_t n, uint32_t* order, uint32_t* child, uint32_t* range,
                  std::string& error);

// Builds tris, tri_of_global and the wide BVH into `out` (bvh_build.hip).
bool build_bvh(const BuildInput& in, SceneBuffers& out, hipStream_t stream, double& gpu_ms,
               std::string& error);

// Re-places every triangle from the (changed) positions and transforms in
// `in`, keeping the BVH's topology and leaf order, and recomputes its boxes
// level by level from the leaves up. The boxes equal those a build of the
// same topology writes (bvh_build.hip).
bool refit_bvh(const BuildInput& in, SceneBuffers& s, hipStream_t stream, double& gpu_ms, std::string& error);

// Rewrites the triangles' covered bits from instance_covered (null: none)
// after the instances' materials changed (bvh_build.hip).
hipError_t mark_covered(SceneBuffers& s, uint32_t tri_count, const uint32_t* instance_covered, hipStream_t stream);

// Test hook of the sort (bvh_build.hip): stable sort of keys with values.
bool radix_sort_pairs(uint32_t* keys, uint32_t* values, uint32_t n, uint32_t key_bits,
                      hipStream_t stream, std::string& error);

}  // namespace lotus
