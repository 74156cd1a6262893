#pragma once
// Source compatibility with scene programs and tests written against the
// reference (include/raymotion/metal.hpp:6-7): the Metal entry point name now
// renders through the HIP backend.
#include "raymotion/hip.hpp"

namespace raymotion {
bool render_image_metal(const Scene& scene, const Camera& camera, const RenderConfig& config, uint64_t seed,
                        std::vector<uint8_t>& output, std::string& error);
}  // namespace raymotion
