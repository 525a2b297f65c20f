// Minimal JSON reader for dataset descriptions (transforms.json). Values are
// null, bool, number (double), string, array or object; objects keep their
// key order. `value(key, default)` mirrors the lookup the reference's loader
// uses: the default when the key is absent, a conversion of the stored number
// when present, an error when the stored value is not a number.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace turbo {

struct JsonValue {
    enum class Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Kind::Null;
    bool boolean = false;
    double number = 0.0;
    std::string string;
    std::vector<JsonValue> array;
    std::vector<std::pair<std::string, JsonValue>> object;

    bool is_number() const { return kind == Kind::Number; }
    bool is_string() const { return kind == Kind::String; }
    bool is_array() const { return kind == Kind::Array; }
    bool is_object() const { return kind == Kind::Object; }

    const JsonValue* find(const std::string& key) const {
        if (kind != Kind::Object) return nullptr;
        for (const auto& kv : object)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool contains(const std::string& key) const { return find(key) != nullptr; }
    const JsonValue& at(const std::string& key) const {
        const JsonValue* v = find(key);
        if (!v) throw std::runtime_error("JSON: missing key \"" + key + "\"");
        return *v;
    }
    const JsonValue& at(size_t i) const {
        if (kind != Kind::Array || i >= array.size()) throw std::runtime_error("JSON: index out of range");
        return array[i];
    }
    size_t size() const { return kind == Kind::Array ? array.size() : kind == Kind::Object ? object.size() : 0; }

    double as_number(const char* what = "value") const {
        if (kind != Kind::Number) throw std::runtime_error(std::string("JSON: ") + what + " is not a number");
        return number;
    }
    float as_float(const char* what = "value") const { return (float)as_number(what); }
    int as_int(const char* what = "value") const { return (int)as_number(what); }
    const std::string& as_string(const char* what = "value") const {
        if (kind != Kind::String) throw std::runtime_error(std::string("JSON: ") + what + " is not a string");
        return string;
    }
    float value(const std::string& key, float dflt) const {
        const JsonValue* v = find(key);
        return v ? v->as_float(key.c_str()) : dflt;
    }
    int value(const std::string& key, int dflt) const {
        const JsonValue* v = find(key);
        return v ? v->as_int(key.c_str()) : dflt;
    }
};

// Parses a complete JSON text; throws std::runtime_error naming the byte
// offset of the first error.
JsonValue parse_json(const std::string& text);

}  // namespace turbo
