#pragma once
// Wavefront OBJ/MTL input (interface of include/raymotion/obj.hpp in the
// reference): triangles and polygon fans, optional UVs and normals, negative
// indices; MTL Kd/Ke/Ni/Pr/Pm and map_Kd. Textures are read from binary or
// ASCII PPM files and decoded from sRGB to linear.
#include <string>

#include "raymotion/scene.hpp"

namespace raymotion {

// Adds the triangles of an OBJ file, placed at (ox, oy, oz) with uniform
// scale and the given default material. Returns the number added.
int add_obj_file(Scene& scene, const std::string& path, double ox, double oy, double oz, double scale, double ar,
                 double ag, double ab, double er, double eg, double eb, double ior, double rough, int shader,
                 double metallic = 0);
// OBJ with materials from an MTL file (textures resolved relative to it).
int load_obj_materials(Scene& scene, const std::string& obj, const std::string& mtl);
// Parses OBJ text with the rules of add_obj_file.
int add_obj_tokens(Scene& scene, const std::string& data, double ox, double oy, double oz, double scale, double ar,
                   double ag, double ab, double er, double eg, double eb, double ior, double rough, int shader,
                   double metallic = 0);

}  // namespace raymotion
