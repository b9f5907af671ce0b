// Image files for scenes and renders: texture decoding (PNG), environment maps (Radiance .hdr, PFM,
// or an 8-bit image read as sRGB) and lossless outputs (PFM, PNG). Replaces the reference's image
// functions (src/scene/image.h:12-22), which use stb_image; the PNG codec here is written over zlib.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "scene.hpp"

namespace phh {

// Decodes a PNG (every colour type and bit depth, interlaced or not) or a baseline JPEG to RGBA8, row 0
// at the top. PNG: 16-bit samples keep their high byte; grey and palette images expand to RGBA, and a
// tRNS chunk supplies alpha. Other formats fail with a message naming the format.
bool decode_image(const uint8_t* data, size_t size, Image8& out, std::string& error);
// Baseline (sequential Huffman, 8-bit) JPEG to RGBA8; decode_image dispatches here (image_jpeg.cpp).
bool decode_jpeg(const uint8_t* data, size_t size, Image8& out, std::string& error);
bool load_image(const std::string& path, Image8& out, std::string& error);

// An equirectangular environment: Radiance RGBE (.hdr, flat or run-length encoded scanlines), PFM, or
// an 8-bit image whose sRGB values are linearised. Non-finite and negative texels become 0.
bool load_environment(const std::string& path, Backdrop& env, std::string& error);

// RGBA8 (R in the low byte), rows top to bottom.
bool save_png(const std::string& path, int width, int height, const uint32_t* rgba);
// Portable float map: "PF", width and height, scale -1 (little endian), rows bottom to top.
bool save_pfm(const std::string& path, int width, int height, const float* rgb);
bool load_pfm(const std::string& path, int& width, int& height, std::vector<float>& rgb);

}  // namespace phh
