// Lotus::OffscreenRenderer over liblotus_hip: the C++ facade a hydra-lotus
// tree compiles (beside backend/vulkan, against include/lotus/*.hpp) so that
// lotus-headless, hdLotus and the viewport drive the HIP path tracer through
// the interface they already use (include/lotus/vulkan_backend.hpp:753-842).
//
// The facade owns one lotus_hip_renderer. UpdateScene keeps the resident
// geometry and textures a SceneUpdate names alive by their shared pointers
// (identity is the address, include/lotus/extraction.hpp:86-90), applies the
// plan's instance, material, environment and light replacements, and hands
// the whole scene to lotus_hip_scene_apply, which compares it with the
// resident one and retains, refits or rebuilds (gpu_scene.cpp:1606-1725).
// RenderScene and RenderTemporalScene map the target and settings onto
// lotus_hip_frame_params and the results back onto GpuFrameEvidence.
#pragma once

#include <memory>
#include <string>

#include <lotus/vulkan_backend.hpp>

namespace Lotus {

// CreateOffscreenRenderer's counterpart (vulkan_backend.hpp:879-883) for the
// HIP backend: null with `status` Skip when no gfx950 device is present,
// Fail when one is present but cannot be used; `error` holds the detail.
// The renderer runs on HIP device `device`.
[[nodiscard]] std::unique_ptr<OffscreenRenderer> CreateHipOffscreenRenderer(FrameStatus& status,
                                                                            std::string& error,
                                                                            int device = 0);

}  // namespace Lotus
