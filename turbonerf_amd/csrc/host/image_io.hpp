// Training-image decoding to RGBA8 and render output as PNG (the reference
// uses stb_image's stbi_load(..., 4), src/models/training-image.cu:11-13, and
// stbi_write_png, src/utils/gpu-image.cu:47-83). PNG: every colour type and
// bit depth of non-interlaced PNG, converted as stb_image converts them --
// grey scaled to 8 bits (x255, x85, x17 for 1-, 2-, 4-bit), 16-bit samples
// reduced to their high byte, palette alpha and grey/RGB colour keys from
// tRNS, alpha 255 otherwise. JPEG (jpeg.cpp): baseline, extended and
// progressive Huffman, 8-bit grey or YCbCr/RGB, any sampling factors,
// restart intervals; alpha 255.
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

// Decodes a JPEG from bytes in memory; throws std::runtime_error naming the
// file on any error (truncated or corrupt data, unsupported coding).
Rgba8Image decode_jpeg_rgba8(const uint8_t* data, size_t size, const std::string& name);

// Width and height from a JPEG's frame header; false when there is none.
bool read_jpeg_size(const uint8_t* data, size_t size, int& width, int& height);

// A PNG or JPEG file, told apart by its signature.
Rgba8Image load_image_rgba8(const std::string& path);
bool read_image_size(const std::string& path, int& width, int& height);

// Writes an 8-bit RGBA PNG (unfiltered rows, zlib level 6); throws on I/O errors.
void write_png_rgba8(const std::string& path, int width, int height, const uint8_t* rgba);

}  // namespace turbo
