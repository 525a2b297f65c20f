// JPEG decoding (ITU-T T.81 / JFIF) to RGBA8 for training images, in place of
// stb_image's JPEG path (reference src/models/training-image.cu:11-13). Own
// implementation from the standard:
//
// * frames: baseline and extended sequential Huffman (SOF0/SOF1) and
//   progressive Huffman (SOF2), 8-bit samples, 1 component (grey) or 3
//   (YCbCr, or RGB when an Adobe APP14 marker says so), sampling factors
//   1..4, restart intervals; arithmetic coding, lossless and hierarchical
//   frames, 12-bit samples and 4-component (CMYK) images are refused;
// * all scans decode into per-component DCT coefficient planes (so the
//   progressive spectral-selection and successive-approximation scans
//   refine the same blocks), then each block is dequantised and inverse
//   transformed (separable float DCT-III, exact to fp32 rounding);
// * chroma planes sampled at half the luma rate are upsampled with the
//   triangle ("fancy") filter that libjpeg and stb_image use (weights 3/4
//   and 1/4 per direction), other ratios by replication; YCbCr goes to RGB
//   with the JFIF equations, rounded to nearest.
//
// Results match libjpeg's integer pipeline (which Pillow uses) to within a
// few levels per channel: the IDCT and colour conversion round differently.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "image_io.hpp"

namespace turbo {

namespace {

// zig-zag order: natural index of the k-th coefficient
constexpr uint8_t kZigZag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

[[noreturn]] void fail(const std::string& name, const std::string& what) {
    throw std::runtime_error(name + ": " + what);
}

// Canonical Huffman table (T.81 Annex C): codes of each length are
// consecutive integers; a 9-bit lookup resolves short codes at once.
struct Huffman {
    bool defined = false;
    uint8_t values[256] = {};
    int32_t maxcode[18] = {};  // largest code of length l, -1 if none
    int32_t valptr[17] = {};   // index into values of the first code of length l
    int32_t mincode[17] = {};
    uint16_t fast[512] = {};   // (length << 8) | value for codes up to 9 bits, 0 = slow path

    void build(const uint8_t* counts, const uint8_t* vals, int n_vals) {
        std::memcpy(values, vals, (size_t)n_vals);
        std::memset(fast, 0, sizeof(fast));
        int32_t code = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            valptr[l] = k;
            mincode[l] = code;
            code += counts[l - 1];
            k += counts[l - 1];
            maxcode[l] = counts[l - 1] ? code - 1 : -1;
            // codes of length l: mincode .. code - 1
            if (l <= 9)
                for (int32_t c = mincode[l]; c < code; ++c) {
                    const int shift = 9 - l;
                    for (int32_t f = c << shift; f < ((c + 1) << shift); ++f)
                        fast[f] = (uint16_t)((l << 8) | values[valptr[l] + (c - mincode[l])]);
                }
            code <<= 1;
        }
        maxcode[17] = 0x7FFFFFFF;
        defined = true;
    }
};

// Entropy-coded segment reader: byte stuffing (FF 00) removed; at a marker
// the stream yields zero bits (the marker is left for the caller).
struct BitReader {
    const uint8_t* p;
    const uint8_t* end;
    uint64_t acc = 0;
    int bits = 0;
    bool at_marker = false;

    void fill() {
        while (bits <= 56) {
            uint32_t byte = 0;
            if (!at_marker && p < end) {
                byte = *p;
                if (byte == 0xFF) {
                    const uint8_t next = p + 1 < end ? p[1] : 0;
                    if (next == 0x00) {
                        p += 2;
                    } else {
                        at_marker = true;  // a marker: zeros from here
                        byte = 0;
                    }
                } else {
                    ++p;
                }
            }
            acc |= (uint64_t)byte << (56 - bits);
            bits += 8;
        }
    }
    uint32_t peek(int n) {
        if (bits < n) fill();
        return (uint32_t)(acc >> (64 - n));
    }
    void skip(int n) {
        acc <<= n;
        bits -= n;
    }
    uint32_t get(int n) {
        if (n == 0) return 0;
        const uint32_t v = peek(n);
        skip(n);
        return v;
    }
    // T.81 F.2.2.1 EXTEND: a magnitude category s value to its signed value
    int32_t receive_extend(int s) {
        if (s == 0) return 0;
        const int32_t v = (int32_t)get(s);
        return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v;
    }
    int decode(const Huffman& h, const std::string& name) {
        const uint32_t f = peek(9);
        if (h.fast[f]) {
            skip(h.fast[f] >> 8);
            return h.fast[f] & 0xFF;
        }
        // T.81 F.2.2.3 DECODE for codes longer than 9 bits
        const uint32_t code16 = peek(16);
        for (int l = 10; l <= 16; ++l) {
            const int32_t code = (int32_t)(code16 >> (16 - l));
            if (h.maxcode[l] >= 0 && code <= h.maxcode[l] && code >= h.mincode[l]) {
                skip(l);
                return h.values[h.valptr[l] + (code - h.mincode[l])];
            }
        }
        fail(name, "corrupt JPEG data (bad Huffman code)");
    }
    // after an RSTn marker: the reader restarts at the next byte
    void restart(const std::string& name) {
        acc = 0;
        bits = 0;
        at_marker = false;
        // skip to the marker
        while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) {
            if (p[0] == 0xFF && p[1] != 0x00 && p[1] != 0xFF) fail(name, "corrupt JPEG data (missing restart marker)");
            ++p;
        }
        if (p + 1 >= end) fail(name, "truncated JPEG data");
        p += 2;
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int bw = 0, bh = 0;          // coefficient blocks per row / column (padded to whole MCUs)
    int w = 0, ht = 0;           // samples actually covered by the image
    std::vector<int16_t> coef;   // bw * bh blocks of 64 (natural order)
    int dc_pred = 0;
    int td = 0, ta = 0;          // Huffman tables of the current scan
};

struct Decoder {
    std::string name;
    const uint8_t* data;
    size_t size;
    size_t pos = 0;
    uint16_t qt[4][64] = {};
    bool qt_defined[4] = {};
    Huffman dc[4], ac[4];
    std::vector<Component> comp;
    int width = 0, height = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false, frame_seen = false;
    int restart_interval = 0;
    int adobe_transform = -1;  // APP14 colour transform, -1 = no Adobe marker
    int eobrun = 0;

    uint8_t byte() {
        if (pos >= size) fail(name, "truncated JPEG data");
        return data[pos++];
    }
    int word() {
        const int a = byte();
        return (a << 8) | byte();
    }

    void read_dqt(int len) {
        const size_t stop = pos + (size_t)len;
        while (pos < stop) {
            const int pq = byte(), t = pq & 15, prec = pq >> 4;
            if (t > 3 || prec > 1) fail(name, "bad JPEG quantisation table");
            for (int k = 0; k < 64; ++k) qt[t][kZigZag[k]] = (uint16_t)(prec ? word() : byte());
            qt_defined[t] = true;
        }
        if (pos != stop) fail(name, "bad JPEG quantisation table length");
    }

    void read_dht(int len) {
        const size_t stop = pos + (size_t)len;
        while (pos < stop) {
            const int tc = byte(), cls = tc >> 4, t = tc & 15;
            if (cls > 1 || t > 3) fail(name, "bad JPEG Huffman table");
            uint8_t counts[16];
            int n = 0;
            for (int l = 0; l < 16; ++l) n += counts[l] = byte();
            if (n > 256) fail(name, "bad JPEG Huffman table");
            uint8_t vals[256];
            for (int k = 0; k < n; ++k) vals[k] = byte();
            (cls ? ac[t] : dc[t]).build(counts, vals, n);
        }
        if (pos != stop) fail(name, "bad JPEG Huffman table length");
    }

    void read_sof(int marker, int len) {
        if (frame_seen) fail(name, "JPEG with several frames");
        const int precision = byte();
        if (precision != 8) fail(name, "JPEG with " + std::to_string(precision) + "-bit samples is not supported");
        height = word();
        width = word();
        const int nc = byte();
        if (width <= 0 || height <= 0) fail(name, "JPEG without a size (DNL) is not supported");
        if (nc != 1 && nc != 3) fail(name, "JPEG with " + std::to_string(nc) + " components is not supported");
        if (len != 6 + 3 * nc) fail(name, "bad JPEG frame header");
        comp.resize(nc);
        for (auto& c : comp) {
            c.id = byte();
            const int hv = byte();
            c.h = hv >> 4;
            c.v = hv & 15;
            c.tq = byte();
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) fail(name, "bad JPEG component");
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        progressive = marker == 0xC2;
        mcux = (width + 8 * hmax - 1) / (8 * hmax);
        mcuy = (height + 8 * vmax - 1) / (8 * vmax);
        for (auto& c : comp) {
            c.bw = mcux * c.h;
            c.bh = mcuy * c.v;
            c.w = (width * c.h + hmax - 1) / hmax;
            c.ht = (height * c.v + vmax - 1) / vmax;
            c.coef.assign((size_t)c.bw * c.bh * 64, 0);
        }
        frame_seen = true;
    }

    int16_t* block(Component& c, int bx, int by) { return c.coef.data() + ((size_t)by * c.bw + bx) * 64; }

    // one block of a sequential scan (T.81 F.2.2)
    void decode_block_seq(BitReader& br, Component& c, int16_t* b) {
        const int t = br.decode(dc[c.td], name);
        if (t > 11) fail(name, "corrupt JPEG data (DC category)");
        c.dc_pred += br.receive_extend(t);
        b[0] = (int16_t)c.dc_pred;
        for (int k = 1; k < 64;) {
            const int rs = br.decode(ac[c.ta], name), r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break;  // end of block
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) fail(name, "corrupt JPEG data (AC index)");
            b[kZigZag[k]] = (int16_t)br.receive_extend(s);
            ++k;
        }
    }

    // progressive scans (T.81 G.1.2)
    void decode_dc_first(BitReader& br, Component& c, int16_t* b, int al) {
        const int t = br.decode(dc[c.td], name);
        if (t > 11) fail(name, "corrupt JPEG data (DC category)");
        c.dc_pred += br.receive_extend(t);
        b[0] = (int16_t)(c.dc_pred * (1 << al));
    }
    void decode_dc_refine(BitReader& br, int16_t* b, int al) {
        if (br.get(1)) b[0] = (int16_t)(b[0] | (1 << al));
    }
    void decode_ac_first(BitReader& br, Component& c, int16_t* b, int ss, int se, int al) {
        if (eobrun > 0) {
            --eobrun;
            return;
        }
        for (int k = ss; k <= se;) {
            const int rs = br.decode(ac[c.ta], name), r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r < 15) {  // EOBr: this block and eobrun more end here
                    eobrun = (1 << r) - 1;
                    if (r) eobrun += (int)br.get(r);
                    break;
                }
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) fail(name, "corrupt JPEG data (AC index)");
            b[kZigZag[k]] = (int16_t)(br.receive_extend(s) * (1 << al));
            ++k;
        }
    }
    // a correction bit for a coefficient already non-zero: adds 2^al away from zero
    static void refine(BitReader& br, int16_t* coef, int p1) {
        if (br.get(1) && (*coef & p1) == 0) *coef = (int16_t)(*coef >= 0 ? *coef + p1 : *coef - p1);
    }
    void decode_ac_refine(BitReader& br, Component& c, int16_t* b, int ss, int se, int al) {
        const int p1 = 1 << al;
        int k = ss;
        if (eobrun == 0) {
            for (; k <= se;) {
                const int rs = br.decode(ac[c.ta], name);
                int r = rs >> 4;
                const int s = rs & 15;
                int val = 0;
                if (s == 0) {
                    if (r < 15) {
                        eobrun = 1 << r;
                        if (r) eobrun += (int)br.get(r);
                        break;  // the rest of the block takes correction bits below
                    }
                    // ZRL: 16 coefficients with zero history are skipped
                } else {
                    if (s != 1) fail(name, "corrupt JPEG data (refinement)");
                    val = br.get(1) ? p1 : -p1;
                }
                // skip r zero-history coefficients (refining the non-zero
                // ones passed on the way), then place val at the next one
                for (; k <= se; ++k) {
                    int16_t* coef = b + kZigZag[k];
                    if (*coef != 0) {
                        refine(br, coef, p1);
                    } else {
                        if (r == 0) {
                            if (val) *coef = (int16_t)val;
                            ++k;
                            break;
                        }
                        --r;
                    }
                }
            }
        }
        if (eobrun > 0) {
            for (; k <= se; ++k) {
                int16_t* coef = b + kZigZag[k];
                if (*coef != 0) refine(br, coef, p1);
            }
            --eobrun;
        }
    }

    void read_sos(int len) {
        if (!frame_seen) fail(name, "JPEG scan before the frame header");
        const int ns = byte();
        if (ns < 1 || ns > 4 || len != 4 + 2 * ns) fail(name, "bad JPEG scan header");
        std::vector<Component*> sc;
        for (int i = 0; i < ns; ++i) {
            const int id = byte(), t = byte();
            Component* c = nullptr;
            for (auto& x : comp)
                if (x.id == id) c = &x;
            if (!c) fail(name, "JPEG scan names an unknown component");
            c->td = t >> 4;
            c->ta = t & 15;
            if (c->td > 3 || c->ta > 3) fail(name, "bad JPEG scan header");
            sc.push_back(c);
        }
        const int ss = byte(), se = byte(), a = byte(), ah = a >> 4, al = a & 15;
        if (progressive) {
            if (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1) || al > 13)
                fail(name, "bad progressive JPEG scan");
        } else if (ss != 0 || se != 63 || a != 0) {
            fail(name, "bad sequential JPEG scan");
        }
        for (Component* c : sc) {
            const bool needs_dc = !progressive || (ss == 0 && ah == 0);
            const bool needs_ac = !progressive || ss > 0;
            if ((needs_dc && !dc[c->td].defined) || (needs_ac && !ac[c->ta].defined))
                fail(name, "JPEG scan uses an undefined Huffman table");
            c->dc_pred = 0;
        }
        eobrun = 0;
        BitReader br{data + pos, data + size};
        auto do_block = [&](Component& c, int bx, int by) {
            int16_t* b = block(c, bx, by);
            if (!progressive) decode_block_seq(br, c, b);
            else if (ss == 0) (ah == 0 ? decode_dc_first(br, c, b, al) : decode_dc_refine(br, b, al));
            else if (ah == 0) decode_ac_first(br, c, b, ss, se, al);
            else decode_ac_refine(br, c, b, ss, se, al);
        };
        auto restart_if_due = [&](int& todo, int unit_index, int n_units) {
            if (restart_interval == 0) return;
            if (--todo == 0 && unit_index + 1 < n_units) {
                br.restart(name);
                for (Component* c : sc) c->dc_pred = 0;
                eobrun = 0;
                todo = restart_interval;
            }
        };
        int todo = restart_interval;
        if (ns == 1) {
            // a non-interleaved scan covers the component's own blocks (T.81 A.2.2)
            Component& c = *sc[0];
            const int nbx = (c.w + 7) / 8, nby = (c.ht + 7) / 8;
            for (int by = 0; by < nby; ++by)
                for (int bx = 0; bx < nbx; ++bx) {
                    do_block(c, bx, by);
                    restart_if_due(todo, by * nbx + bx, nbx * nby);
                }
        } else {
            for (int my = 0; my < mcuy; ++my)
                for (int mx = 0; mx < mcux; ++mx) {
                    for (Component* c : sc)
                        for (int v = 0; v < c->v; ++v)
                            for (int h = 0; h < c->h; ++h) do_block(*c, mx * c->h + h, my * c->v + v);
                    restart_if_due(todo, my * mcux + mx, mcux * mcuy);
                }
        }
        // continue after the entropy-coded data (at the next marker)
        pos = (size_t)(br.p - data);
        while (pos + 1 < size && !(data[pos] == 0xFF && data[pos + 1] != 0x00 && !(data[pos + 1] >= 0xD0 && data[pos + 1] <= 0xD7)))
            ++pos;
    }

    void parse() {
        if (size < 4 || data[0] != 0xFF || data[1] != 0xD8) fail(name, "not a JPEG file");
        pos = 2;
        bool scanned = false;
        while (true) {
            // markers may be preceded by fill bytes 0xFF
            uint8_t b = byte();
            if (b != 0xFF) fail(name, "corrupt JPEG data (expected a marker)");
            do b = byte();
            while (b == 0xFF);
            const int m = b;
            if (m == 0xD9) break;  // EOI
            if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;  // no length
            const int len = word() - 2;
            if (len < 0 || pos + (size_t)len > size) fail(name, "truncated JPEG data");
            switch (m) {
            case 0xC0: case 0xC1: case 0xC2: read_sof(m, len); break;
            case 0xC3: case 0xC5: case 0xC6: case 0xC7: case 0xC9: case 0xCA: case 0xCB: case 0xCD: case 0xCE:
            case 0xCF:
                fail(name, "lossless, hierarchical and arithmetic-coded JPEG are not supported");
            case 0xC4: read_dht(len); break;
            case 0xDB: read_dqt(len); break;
            case 0xDD:
                if (len != 2) fail(name, "bad JPEG restart interval");
                restart_interval = word();
                break;
            case 0xDA: read_sos(len); scanned = true; break;
            case 0xEE:  // APP14: "Adobe" and the colour transform flag
                if (len >= 12 && std::memcmp(data + pos, "Adobe", 5) == 0) adobe_transform = data[pos + 11];
                pos += (size_t)len;
                break;
            default: pos += (size_t)len; break;  // APPn, COM, DNL ...
            }
            if (pos >= size) break;  // tolerate a missing EOI
        }
        if (!frame_seen || !scanned) fail(name, "JPEG without image data");
        for (auto& c : comp)
            if (!qt_defined[c.tq]) fail(name, "JPEG component uses an undefined quantisation table");
    }
};

// 8x8 inverse DCT (T.81 A.3.3) of dequantised coefficients, separable, +128,
// rounded and clamped to 0..255
struct Idct {
    float c[8][8];  // c[x][u] = C(u)/2 cos((2x+1) u pi / 16)
    Idct() {
        const double pi = 3.14159265358979323846;
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                c[x][u] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * pi / 16.0));
    }
    void run(const int16_t* in, const uint16_t* q, uint8_t* out, int stride) const {
        float f[64], t[64];
        for (int k = 0; k < 64; ++k) f[k] = (float)in[k] * (float)q[k];
        for (int v = 0; v < 8; ++v)  // rows: t[v][x] = sum_u c[x][u] f[v][u]
            for (int x = 0; x < 8; ++x) {
                float s = 0.0f;
                for (int u = 0; u < 8; ++u) s += c[x][u] * f[v * 8 + u];
                t[v * 8 + x] = s;
            }
        for (int y = 0; y < 8; ++y)  // columns
            for (int x = 0; x < 8; ++x) {
                float s = 0.0f;
                for (int v = 0; v < 8; ++v) s += c[y][v] * t[v * 8 + x];
                const float r = std::nearbyint(s + 128.0f);
                out[y * stride + x] = (uint8_t)(r < 0.0f ? 0.0f : r > 255.0f ? 255.0f : r);
            }
    }
};

// component plane (w x h valid samples, row stride `stride`) upsampled by
// (fx, fy) to the image grid: factor 2 by the triangle filter, others by
// replication
std::vector<uint8_t> upsample(const std::vector<uint8_t>& plane, int stride, int w, int h, int fx, int fy, int W,
                              int H) {
    auto at = [&](int x, int y) -> int {
        x = x < 0 ? 0 : x >= w ? w - 1 : x;
        y = y < 0 ? 0 : y >= h ? h - 1 : y;
        return plane[(size_t)y * stride + x];
    };
    std::vector<uint8_t> out((size_t)W * H);
    for (int Y = 0; Y < H; ++Y) {
        const int y = Y / fy;
        // the nearer and the farther input row for this output row
        const int y_far = fy == 2 ? (Y % 2 == 0 ? y - 1 : y + 1) : y;
        for (int X = 0; X < W; ++X) {
            const int x = X / fx;
            const int x_far = fx == 2 ? (X % 2 == 0 ? x - 1 : x + 1) : x;
            int v;
            if (fx == 2 && fy == 2) {
                const int near_col = 3 * at(x, y) + at(x, y_far), far_col = 3 * at(x_far, y) + at(x_far, y_far);
                v = (3 * near_col + far_col + 8) >> 4;
            } else if (fx == 2 && fy == 1) {
                v = (3 * at(x, y) + at(x_far, y) + 2) >> 2;
            } else if (fx == 1 && fy == 2) {
                v = (3 * at(x, y) + at(x, y_far) + 2) >> 2;
            } else {
                v = at(x, y);
            }
            out[(size_t)Y * W + X] = (uint8_t)v;
        }
    }
    return out;
}

}  // namespace

bool read_jpeg_size(const uint8_t* f, size_t n, int& width, int& height) {
    if (n < 4 || f[0] != 0xFF || f[1] != 0xD8) return false;
    size_t p = 2;
    while (p + 4 <= n) {
        if (f[p] != 0xFF) return false;
        while (p < n && f[p] == 0xFF) ++p;
        if (p >= n) return false;
        const int m = f[p++];
        if (m == 0xD9 || m == 0xDA) return false;
        if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
        if (p + 2 > n) return false;
        const int len = (f[p] << 8) | f[p + 1];
        if ((m >= 0xC0 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
            if (p + 7 > n) return false;
            height = (f[p + 3] << 8) | f[p + 4];
            width = (f[p + 5] << 8) | f[p + 6];
            return width > 0 && height > 0;
        }
        p += (size_t)len;
    }
    return false;
}

Rgba8Image decode_jpeg_rgba8(const uint8_t* data, size_t size, const std::string& name) {
    Decoder d;
    d.name = name;
    d.data = data;
    d.size = size;
    d.parse();
    static const Idct idct;
    const int W = d.width, H = d.height, nc = (int)d.comp.size();
    std::vector<std::vector<uint8_t>> planes(nc);
    for (int i = 0; i < nc; ++i) {
        Component& c = d.comp[i];
        const int stride = c.bw * 8;
        std::vector<uint8_t> plane((size_t)stride * c.bh * 8);
        for (int by = 0; by < c.bh; ++by)
            for (int bx = 0; bx < c.bw; ++bx)
                idct.run(d.block(c, bx, by), d.qt[c.tq], plane.data() + (size_t)by * 8 * stride + bx * 8, stride);
        const int fx = d.hmax / c.h, fy = d.vmax / c.v;
        if (d.hmax % c.h || d.vmax % c.v) fail(name, "JPEG with non-integral sampling ratios is not supported");
        planes[i] = upsample(plane, stride, c.w, c.ht, fx, fy, W, H);
    }
    Rgba8Image img;
    img.width = W;
    img.height = H;
    img.rgba.resize((size_t)W * H * 4);
    // three components are YCbCr unless an Adobe marker says RGB (transform 0)
    const bool ycc = nc == 3 && d.adobe_transform != 0;
    auto clamp8 = [](float v) -> uint8_t {
        const float r = std::nearbyint(v);
        return (uint8_t)(r < 0.0f ? 0.0f : r > 255.0f ? 255.0f : r);
    };
    for (size_t k = 0; k < (size_t)W * H; ++k) {
        uint8_t* o = img.rgba.data() + 4 * k;
        if (nc == 1) {
            o[0] = o[1] = o[2] = planes[0][k];
        } else if (ycc) {
            const float y = planes[0][k], cb = (float)planes[1][k] - 128.0f, cr = (float)planes[2][k] - 128.0f;
            o[0] = clamp8(y + 1.402f * cr);
            o[1] = clamp8(y - 0.344136f * cb - 0.714136f * cr);
            o[2] = clamp8(y + 1.772f * cb);
        } else {
            o[0] = planes[0][k];
            o[1] = planes[1][k];
            o[2] = planes[2][k];
        }
        o[3] = 255;
    }
    return img;
}

}  // namespace turbo
