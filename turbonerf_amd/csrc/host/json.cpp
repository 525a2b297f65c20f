// Recursive-descent JSON reader (RFC 8259 grammar; numbers through from_chars).
#include "json.hpp"

#include <charconv>
#include <cstdlib>
#include <cstring>

namespace turbo {
namespace {

struct Reader {
    const std::string& s;
    size_t i = 0;
    int depth = 0;

    [[noreturn]] void fail(const char* msg) const {
        throw std::runtime_error(std::string("JSON: ") + msg + " at byte " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail((std::string("expected '") + c + "'").c_str());
    }
    bool literal(const char* word) {
        const size_t n = std::strlen(word);
        if (s.compare(i, n, word) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    static void put_utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (i + 4 > s.size()) fail("truncated \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail("bad hex digit in \\u escape");
        }
        return v;
    }
    std::string str() {
        expect('"');
        std::string out;
        while (true) {
            if (i >= s.size()) fail("unterminated string");
            const char c = s[i++];
            if (c == '"') return out;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i >= s.size()) fail("unterminated escape");
            const char e = s[i++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00) {  // high surrogate: a low one must follow
                        if (!(i + 1 < s.size() && s[i] == '\\' && s[i + 1] == 'u')) fail("lone surrogate");
                        i += 2;
                        const uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) fail("bad low surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        fail("lone surrogate");
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
    double num() {
        // validate the RFC 8259 number grammar, then convert with from_chars
        const size_t b = i;
        if (i < s.size() && s[i] == '-') ++i;
        if (i < s.size() && s[i] == '0') {
            ++i;
        } else if (i < s.size() && s[i] >= '1' && s[i] <= '9') {
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        } else {
            fail("bad number");
        }
        if (i < s.size() && s[i] == '.') {
            ++i;
            if (!(i < s.size() && s[i] >= '0' && s[i] <= '9')) fail("bad fraction");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        }
        if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
            ++i;
            if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
            if (!(i < s.size() && s[i] >= '0' && s[i] <= '9')) fail("bad exponent");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        }
        double v = 0.0;
        const auto r = std::from_chars(s.data() + b, s.data() + i, v);  // locale-independent
        if (r.ec == std::errc::result_out_of_range) {
            // out of double's range: the nearest representable value, as strtod gives
            v = std::strtod(s.substr(b, i - b).c_str(), nullptr);
        } else if (r.ec != std::errc()) {
            fail("bad number");
        }
        return v;
    }
    JsonValue value() {
        if (++depth > 256) fail("nesting too deep");
        ws();
        if (i >= s.size()) fail("unexpected end");
        JsonValue v;
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.kind = JsonValue::Kind::Object;
            if (!eat('}')) {
                do {
                    ws();
                    std::string k = str();
                    expect(':');
                    v.object.emplace_back(std::move(k), value());
                } while (eat(','));
                expect('}');
            }
        } else if (c == '[') {
            ++i;
            v.kind = JsonValue::Kind::Array;
            if (!eat(']')) {
                do {
                    v.array.push_back(value());
                } while (eat(','));
                expect(']');
            }
        } else if (c == '"') {
            v.kind = JsonValue::Kind::String;
            v.string = str();
        } else if (literal("true")) {
            v.kind = JsonValue::Kind::Bool;
            v.boolean = true;
        } else if (literal("false")) {
            v.kind = JsonValue::Kind::Bool;
        } else if (literal("null")) {
            v.kind = JsonValue::Kind::Null;
        } else {
            v.kind = JsonValue::Kind::Number;
            v.number = num();
        }
        --depth;
        return v;
    }
};

}  // namespace

JsonValue parse_json(const std::string& text) {
    Reader r{text};
    JsonValue v = r.value();
    r.ws();
    if (r.i != text.size()) r.fail("trailing characters");
    return v;
}

}  // namespace turbo
