#pragma once
// C++ entry point of the HIP backend with the signature of the reference's
// render_image_metal (include/raymotion/metal.hpp:6-7).
#include <cstdint>
#include <string>
#include <vector>

#include "raymotion/pathtrace.hpp"

namespace raymotion {
// Renders one frame on the GPU into `output` (W*H*3 RGB8, top row first).
// Returns false with a diagnostic in `error` instead of throwing. Reads the
// render switches RAYMOTION_RESTIR, RAYMOTION_DENOISE, RAYMOTION_RESTIR_TEMPORAL,
// RAYMOTION_RESTIR_CANDIDATES, RAYMOTION_RESTIR_SPATIAL, UOW2_NOADAPT and
// UOW2_NOMIS like the reference, plus RAYMOTION_HIP_DEVICE (device index).
// config.gpu_tree > 0 has the tree built (2) or refit (1) on the GPU instead of
// taken from scene.nodes; the scene runtime sets it unless RAYMOTION_GPU_BVH=0.
bool render_image_hip(const Scene& scene, const Camera& camera, const RenderConfig& config, uint64_t seed,
                      std::vector<uint8_t>& output, std::string& error);
// Same, also returning the linear image that was tonemapped (W*H*3 floats).
bool render_image_hip_hdr(const Scene& scene, const Camera& camera, const RenderConfig& config, uint64_t seed,
                          std::vector<uint8_t>& output, std::vector<float>* hdr, std::string& error);
}  // namespace raymotion
