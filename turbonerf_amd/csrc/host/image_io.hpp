// Training-image decoding to RGBA8 and render output as PNG (the reference
// uses stb_image's stbi_load(..., 4), src/models/training-image.cu:11-13, and
// stbi_write_png, src/utils/gpu-image.cu:47-83). Decoding: every
// colour type and bit depth of non-interlaced PNG, converted as stb_image
// converts them -- grey scaled to 8 bits (x255, x85, x17 for 1-, 2-, 4-bit),
// 16-bit samples reduced to their high byte, palette alpha and grey/RGB colour
// keys from tRNS, alpha 255 otherwise.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace turbo {

struct Rgba8Image {
    int width = 0;
    int height = 0;
    std::vector<uint8_t> rgba;  // width * height * 4, row-major from the top row
};

// Width and height from a PNG's header; false when the file is not a PNG.
bool read_png_size(const std::string& path, int& width, int& height);

// Decodes a PNG file; throws std::runtime_error naming the file on any error
// (unreadable, not a PNG, JPEG, interlaced, corrupt data, bad CRC).
Rgba8Image load_png_rgba8(const std::string& path);

// Same, from bytes in memory; `name` is used in error messages.
Rgba8Image decode_png_rgba8(const uint8_t* data, size_t size, const std::string& name);

// Writes an 8-bit RGBA PNG (unfiltered rows, zlib level 6); throws on I/O errors.
void write_png_rgba8(const std::string& path, int width, int height, const uint8_t* rgba);

}  // namespace turbo
