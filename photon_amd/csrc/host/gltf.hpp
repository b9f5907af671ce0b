// glTF 2.0 scene files: import of .gltf (external or data-URI buffers) and binary .glb into a host
// Scene, and export of a scene's meshes, materials, camera and punctual lights as a .gltf with one
// embedded buffer. Replaces load_gltf / save_gltf (src/scene/gltf.h:12-19).
#pragma once

#include <string>

#include "scene.hpp"

namespace phh {

// Appends the default scene's triangle meshes (lists, strips and fans, under the node hierarchy),
// PBR metallic-roughness materials with KHR_materials_transmission / _ior / _emissive_strength,
// textures with samplers, perspective cameras and KHR_lights_punctual lights. On failure returns false
// with a message; what was appended before the failure stays in the scene.
bool load_gltf(const std::string& path, Scene& scene, std::string& error);

// Writes meshes (one primitive per material), materials, the camera and punctual lights. Textures and
// the environment are not written.
bool save_gltf(const std::string& path, const Scene& scene, std::string& error);

}  // namespace phh
