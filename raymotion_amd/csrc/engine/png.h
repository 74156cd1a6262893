// PNG decoding for OBJ/MTL textures (map_Kd). The reference loads textures
// with stb_image into 8-bit RGB (src/engine/obj.cpp:118-124); this decoder
// gives the same bytes for the PNGs that reach it: every colour type and bit
// depth of the PNG specification, both interlace methods, the five row
// filters, and its own inflate (RFC 1950/1951). 16-bit samples keep their high
// byte, grey levels below 8 bits are scaled to 0..255, grey and palette
// images expand to RGB, and alpha is dropped.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace raymotion {

struct Rgb8Image {
    int width = 0, height = 0;
    std::vector<uint8_t> rgb;  // width * height * 3, rows top to bottom
};

bool is_png(const std::string& bytes);

// Throws std::runtime_error with a reason on malformed or unsupported input.
Rgb8Image decode_png(const std::string& bytes);

// zlib stream (RFC 1950) to at most max_out bytes; throws on malformed input,
// a bad Adler-32, or more output than max_out.
std::vector<uint8_t> zlib_inflate(const uint8_t* data, size_t size, size_t max_out);

}  // namespace raymotion
