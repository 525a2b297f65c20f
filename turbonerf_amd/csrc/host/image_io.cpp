// PNG decoding (ISO/IEC 15948): chunk walk with CRC checks, zlib inflate of
// the IDAT stream, per-row unfiltering (None, Sub, Up, Average, Paeth) and
// expansion to RGBA8.
#include "image_io.hpp"

#include <zlib.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace turbo {
namespace {

const uint8_t kSignature[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};

uint32_t be32(const uint8_t* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}
uint32_t be16(const uint8_t* p) { return ((uint32_t)p[0] << 8) | (uint32_t)p[1]; }

std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open image: " + path);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

int channels_of(int color_type) {
    switch (color_type) {
        case 0: return 1;  // grey
        case 2: return 3;  // RGB
        case 3: return 1;  // palette index
        case 4: return 2;  // grey + alpha
        case 6: return 4;  // RGBA
        default: return 0;
    }
}

bool depth_allowed(int color_type, int depth) {
    switch (color_type) {
        case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
        case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
        case 2:
        case 4:
        case 6: return depth == 8 || depth == 16;
        default: return false;
    }
}

uint8_t paeth(uint8_t a, uint8_t b, uint8_t c) {
    const int p = (int)a + (int)b - (int)c;
    const int pa = std::abs(p - (int)a), pb = std::abs(p - (int)b), pc = std::abs(p - (int)c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

}  // namespace

bool read_png_size(const std::string& path, int& width, int& height) {
    std::ifstream in(path, std::ios::binary);
    uint8_t head[24];
    if (!in || !in.read(reinterpret_cast<char*>(head), sizeof(head))) return false;
    if (std::memcmp(head, kSignature, 8) != 0 || std::memcmp(head + 12, "IHDR", 4) != 0) return false;
    width = (int)be32(head + 16);
    height = (int)be32(head + 20);
    return width > 0 && height > 0;
}

Rgba8Image decode_png_rgba8(const uint8_t* f, size_t n, const std::string& name) {
    auto fail = [&](const std::string& msg) -> void { throw std::runtime_error("PNG " + name + ": " + msg); };
    if (n >= 2 && f[0] == 0xFF && f[1] == 0xD8) fail("this is a JPEG file (decode it with decode_jpeg_rgba8)");
    if (n < 8 || std::memcmp(f, kSignature, 8) != 0) fail("not a PNG file");

    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1;
    bool have_ihdr = false, have_iend = false;
    uint8_t pal[256][4];
    int n_pal = 0;
    bool have_key = false;
    uint32_t key[3] = {0, 0, 0};
    std::vector<uint8_t> idat;
    size_t p = 8;
    while (!have_iend) {
        if (p + 12 > n) fail("truncated chunk");
        const uint32_t len = be32(f + p);
        if (len > n - p - 12) fail("truncated chunk");
        const uint8_t* type = f + p + 4;
        const uint8_t* data = f + p + 8;
        const uint32_t crc = be32(data + len);
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc) fail("chunk CRC mismatch");
        const bool first = p == 8;
        if (first != (std::memcmp(type, "IHDR", 4) == 0)) fail("IHDR must come first");
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (len != 13) fail("bad IHDR");
            w = be32(data);
            h = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            if (w == 0 || h == 0 || w > (1u << 24) || h > (1u << 24)) fail("bad image size");
            if ((uint64_t)w * h > (1ull << 30)) fail("image larger than 2^30 pixels");
            if (!depth_allowed(ctype, depth)) fail("unsupported colour type / bit depth");
            if (data[10] != 0 || data[11] != 0) fail("unknown compression or filter method");
            if (data[12] != 0) fail("interlaced PNG is not supported");
            have_ihdr = true;
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            if (len % 3 != 0 || len / 3 > 256 || len == 0) fail("bad PLTE");
            n_pal = (int)(len / 3);
            for (int k = 0; k < n_pal; ++k) {
                pal[k][0] = data[3 * k];
                pal[k][1] = data[3 * k + 1];
                pal[k][2] = data[3 * k + 2];
                pal[k][3] = 255;
            }
        } else if (std::memcmp(type, "tRNS", 4) == 0) {
            if (ctype == 3) {
                if (n_pal == 0 || len > (uint32_t)n_pal) fail("bad tRNS");
                for (uint32_t k = 0; k < len; ++k) pal[k][3] = data[k];
            } else if (ctype == 0) {
                if (len != 2) fail("bad tRNS");
                key[0] = be16(data);
                have_key = true;
            } else if (ctype == 2) {
                if (len != 6) fail("bad tRNS");
                for (int k = 0; k < 3; ++k) key[k] = be16(data + 2 * k);
                have_key = true;
            } else {
                fail("tRNS on an image with alpha");
            }
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), data, data + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            have_iend = true;
        } else if (!(type[0] & 0x20)) {
            fail(std::string("unknown critical chunk ") + std::string(reinterpret_cast<const char*>(type), 4));
        }
        p += 12 + (size_t)len;
    }
    if (!have_ihdr || idat.empty()) fail("missing IHDR or IDAT");
    if (ctype == 3 && n_pal == 0) fail("palette image without PLTE");

    const int ch = channels_of(ctype);
    const size_t row_bytes = ((size_t)w * ch * depth + 7) / 8;
    const size_t bpp = std::max<size_t>(1, (size_t)ch * depth / 8);
    const size_t raw_size = (size_t)h * (row_bytes + 1);
    std::vector<uint8_t> raw(raw_size);
    uLongf got = (uLongf)raw_size;
    const int zr = uncompress(raw.data(), &got, idat.data(), (uLong)idat.size());
    if (zr != Z_OK || got != raw_size) fail("corrupt image data");

    // unfilter in place: row y's bytes follow its filter-type byte
    std::vector<uint8_t> zero(row_bytes, 0);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t* row = raw.data() + (size_t)y * (row_bytes + 1);
        const uint8_t ft = row[0];
        uint8_t* cur = row + 1;
        const uint8_t* up = y ? raw.data() + (size_t)(y - 1) * (row_bytes + 1) + 1 : zero.data();
        switch (ft) {
            case 0: break;
            case 1:
                for (size_t i = bpp; i < row_bytes; ++i) cur[i] = (uint8_t)(cur[i] + cur[i - bpp]);
                break;
            case 2:
                for (size_t i = 0; i < row_bytes; ++i) cur[i] = (uint8_t)(cur[i] + up[i]);
                break;
            case 3:
                for (size_t i = 0; i < row_bytes; ++i) {
                    const unsigned left = i >= bpp ? cur[i - bpp] : 0u;
                    cur[i] = (uint8_t)(cur[i] + ((left + up[i]) >> 1));
                }
                break;
            case 4:
                for (size_t i = 0; i < row_bytes; ++i) {
                    const uint8_t a = i >= bpp ? cur[i - bpp] : 0, c = i >= bpp ? up[i - bpp] : 0;
                    cur[i] = (uint8_t)(cur[i] + paeth(a, up[i], c));
                }
                break;
            default: fail("bad row filter type");
        }
    }

    // expand to RGBA8
    static const uint8_t grey_scale[9] = {0, 0xFF, 0x55, 0, 0x11, 0, 0, 0, 0x01};
    const uint32_t mask = depth < 8 ? (1u << depth) - 1u : 0u;
    auto sample = [&](const uint8_t* row, size_t idx) -> uint32_t {
        if (depth == 8) return row[idx];
        if (depth == 16) return be16(row + 2 * idx);
        const size_t bit = idx * (size_t)depth;
        return (uint32_t)(row[bit >> 3] >> (8 - depth - (int)(bit & 7))) & mask;
    };
    auto to8 = [&](uint32_t v) -> uint8_t { return depth == 16 ? (uint8_t)(v >> 8) : (uint8_t)(v * grey_scale[depth < 8 ? depth : 8]); };
    Rgba8Image img;
    img.width = (int)w;
    img.height = (int)h;
    img.rgba.resize((size_t)w * h * 4);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* row = raw.data() + (size_t)y * (row_bytes + 1) + 1;
        uint8_t* out = img.rgba.data() + (size_t)y * w * 4;
        for (uint32_t x = 0; x < w; ++x, out += 4) {
            if (ctype == 3) {
                const uint32_t idx = sample(row, x);
                if (idx >= (uint32_t)n_pal) fail("palette index out of range");
                std::memcpy(out, pal[idx], 4);
            } else if (ctype == 0) {
                const uint32_t v = sample(row, x);
                const uint8_t g = to8(v);
                out[0] = out[1] = out[2] = g;
                bool clear;
                if (depth == 16) clear = have_key && v == key[0];
                else clear = have_key && g == (uint8_t)((key[0] & 255u) * grey_scale[depth]);
                out[3] = clear ? 0 : 255;
            } else if (ctype == 4) {
                out[0] = out[1] = out[2] = to8(sample(row, 2 * (size_t)x));
                out[3] = to8(sample(row, 2 * (size_t)x + 1));
            } else if (ctype == 2) {
                uint32_t v[3];
                for (int c = 0; c < 3; ++c) {
                    v[c] = sample(row, 3 * (size_t)x + c);
                    out[c] = to8(v[c]);
                }
                bool clear;
                if (depth == 16) clear = have_key && v[0] == key[0] && v[1] == key[1] && v[2] == key[2];
                else clear = have_key && v[0] == (key[0] & 255u) && v[1] == (key[1] & 255u) && v[2] == (key[2] & 255u);
                out[3] = clear ? 0 : 255;
            } else {  // 6
                for (int c = 0; c < 4; ++c) out[c] = to8(sample(row, 4 * (size_t)x + c));
            }
        }
    }
    return img;
}

void write_png_rgba8(const std::string& path, int width, int height, const uint8_t* rgba) {
    if (width <= 0 || height <= 0 || !rgba) throw std::runtime_error("write_png: empty image " + path);
    const size_t row = (size_t)width * 4;
    std::vector<uint8_t> raw((row + 1) * (size_t)height);
    for (int y = 0; y < height; ++y) {
        raw[(size_t)y * (row + 1)] = 0;  // filter type None
        std::memcpy(raw.data() + (size_t)y * (row + 1) + 1, rgba + (size_t)y * row, row);
    }
    uLongf zsize = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zsize);
    if (compress2(z.data(), &zsize, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        throw std::runtime_error("write_png: compression failed for " + path);
    std::vector<uint8_t> out(kSignature, kSignature + 8);
    auto chunk = [&out](const char* type, const uint8_t* data, size_t len) {
        const uint8_t l[4] = {(uint8_t)(len >> 24), (uint8_t)(len >> 16), (uint8_t)(len >> 8), (uint8_t)len};
        out.insert(out.end(), l, l + 4);
        const size_t t0 = out.size();
        out.insert(out.end(), type, type + 4);
        out.insert(out.end(), data, data + len);
        const uint32_t crc = (uint32_t)crc32(crc32(0L, Z_NULL, 0), out.data() + t0, (uInt)(len + 4));
        const uint8_t c[4] = {(uint8_t)(crc >> 24), (uint8_t)(crc >> 16), (uint8_t)(crc >> 8), (uint8_t)crc};
        out.insert(out.end(), c, c + 4);
    };
    const uint8_t ihdr[13] = {(uint8_t)(width >> 24), (uint8_t)(width >> 16), (uint8_t)(width >> 8), (uint8_t)width,
                              (uint8_t)(height >> 24), (uint8_t)(height >> 16), (uint8_t)(height >> 8),
                              (uint8_t)height, 8, 6, 0, 0, 0};
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), zsize);
    chunk("IEND", nullptr, 0);
    std::ofstream f(path, std::ios::binary);
    if (!f || !f.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)out.size()))
        throw std::runtime_error("write_png: cannot write " + path);
}

Rgba8Image load_png_rgba8(const std::string& path) {
    const std::vector<uint8_t> bytes = read_file(path);
    return decode_png_rgba8(bytes.data(), bytes.size(), path);
}

Rgba8Image load_image_rgba8(const std::string& path) {
    const std::vector<uint8_t> bytes = read_file(path);
    if (bytes.size() >= 2 && bytes[0] == 0xFF && bytes[1] == 0xD8)
        return decode_jpeg_rgba8(bytes.data(), bytes.size(), "JPEG " + path);
    return decode_png_rgba8(bytes.data(), bytes.size(), path);
}

bool read_image_size(const std::string& path, int& width, int& height) {
    if (read_png_size(path, width, height)) return true;
    // a JPEG's frame header may follow large APPn segments: read the file
    std::ifstream in(path, std::ios::binary);
    uint8_t head[2];
    if (!in || !in.read(reinterpret_cast<char*>(head), 2) || head[0] != 0xFF || head[1] != 0xD8) return false;
    const std::vector<uint8_t> bytes = read_file(path);
    return read_jpeg_size(bytes.data(), bytes.size(), width, height);
}

}  // namespace turbo
