// Scene front end: JSON scene files and .fpt presets into an FptRenderConfig,
// and the typed-program passes that run on it before a render: the basic
// optimiser, the geometry/shading split, the canonical affine IR, flat-union
// and typed structure-of-arrays lowering (src/scene.rs:743-1976).  Host code
// only; no device work.
//
// Float arithmetic follows the reference's order in single precision with no
// contraction (pragma below), so folded transforms, affine rows and
// normalised plane normals are bit-identical to it.  JSON numbers are parsed
// as doubles (integers through their integer value) and rounded to float
// where the reference does.
#pragma clang fp contract(off)

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <utility>
#include <vector>

#include "fpt_internal.h"

namespace fpt {
namespace {

void seterr(char *error, size_t error_len, const char *fmt, ...) {
    if (!error || !error_len) return;
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(error, error_len, fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------------------
// JSON document (RFC 8259): objects keep their members in order; a repeated
// key resolves to its last occurrence.
// ---------------------------------------------------------------------------
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> items;
    std::vector<std::pair<std::string, Json>> members;

    const Json *get(const char *key) const {
        if (kind != Object) return nullptr;
        const Json *found = nullptr;
        for (const auto &m : members)
            if (m.first == key) found = &m.second;
        return found;
    }
    const double *number() const { return kind == Number ? &num : nullptr; }
    const char *string() const { return kind == String ? str.c_str() : nullptr; }
    bool is_object() const { return kind == Object; }
    bool is_array() const { return kind == Array; }
};

class JsonParser {
  public:
    explicit JsonParser(const std::string &text) : s_(text) {}

    bool parse(Json &out, std::string &err) {
        skip();
        if (!value(out, 0)) {
            err = err_.empty() ? "invalid JSON" : err_;
            return false;
        }
        skip();
        if (i_ != s_.size()) {
            err = "trailing characters after JSON value at offset " + std::to_string(i_);
            return false;
        }
        return true;
    }

  private:
    const std::string &s_;
    size_t i_ = 0;
    std::string err_;

    bool fail(const char *what) {
        if (err_.empty()) err_ = std::string(what) + " at offset " + std::to_string(i_);
        return false;
    }
    void skip() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }
    bool literal(const char *word) {
        const size_t n = std::strlen(word);
        if (s_.compare(i_, n, word) != 0) return false;
        i_ += n;
        return true;
    }
    bool value(Json &v, int depth) {
        if (depth > 256) return fail("JSON nested too deeply");
        if (i_ >= s_.size()) return fail("unexpected end of JSON");
        const char c = s_[i_];
        if (c == '{') return object(v, depth);
        if (c == '[') return array(v, depth);
        if (c == '"') {
            v.kind = Json::String;
            return string(v.str);
        }
        if (c == 't' && literal("true")) {
            v.kind = Json::Bool;
            v.b = true;
            return true;
        }
        if (c == 'f' && literal("false")) {
            v.kind = Json::Bool;
            v.b = false;
            return true;
        }
        if (c == 'n' && literal("null")) {
            v.kind = Json::Null;
            return true;
        }
        if (c == '-' || (c >= '0' && c <= '9')) return number(v);
        return fail("unexpected character in JSON");
    }
    bool object(Json &v, int depth) {
        v.kind = Json::Object;
        ++i_;
        skip();
        if (i_ < s_.size() && s_[i_] == '}') {
            ++i_;
            return true;
        }
        while (true) {
            skip();
            if (i_ >= s_.size() || s_[i_] != '"') return fail("expected object key");
            std::string key;
            if (!string(key)) return false;
            skip();
            if (i_ >= s_.size() || s_[i_] != ':') return fail("expected ':'");
            ++i_;
            skip();
            Json item;
            if (!value(item, depth + 1)) return false;
            v.members.emplace_back(std::move(key), std::move(item));
            skip();
            if (i_ < s_.size() && s_[i_] == ',') {
                ++i_;
                continue;
            }
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return true;
            }
            return fail("expected ',' or '}'");
        }
    }
    bool array(Json &v, int depth) {
        v.kind = Json::Array;
        ++i_;
        skip();
        if (i_ < s_.size() && s_[i_] == ']') {
            ++i_;
            return true;
        }
        while (true) {
            skip();
            Json item;
            if (!value(item, depth + 1)) return false;
            v.items.push_back(std::move(item));
            skip();
            if (i_ < s_.size() && s_[i_] == ',') {
                ++i_;
                continue;
            }
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return true;
            }
            return fail("expected ',' or ']'");
        }
    }
    static void utf8(std::string &out, uint32_t cp) {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xc0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3f));
        } else if (cp < 0x10000) {
            out += char(0xe0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3f));
            out += char(0x80 | (cp & 0x3f));
        } else {
            out += char(0xf0 | (cp >> 18));
            out += char(0x80 | ((cp >> 12) & 0x3f));
            out += char(0x80 | ((cp >> 6) & 0x3f));
            out += char(0x80 | (cp & 0x3f));
        }
    }
    bool hex4(uint32_t &cp) {
        if (i_ + 4 > s_.size()) return fail("truncated \\u escape");
        cp = 0;
        for (int k = 0; k < 4; ++k) {
            const char h = s_[i_++];
            cp <<= 4;
            if (h >= '0' && h <= '9') cp |= uint32_t(h - '0');
            else if (h >= 'a' && h <= 'f') cp |= uint32_t(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') cp |= uint32_t(h - 'A' + 10);
            else return fail("invalid \\u escape");
        }
        return true;
    }
    bool string(std::string &out) {
        ++i_;  // opening quote
        while (true) {
            if (i_ >= s_.size()) return fail("unterminated string");
            const char c = s_[i_++];
            if (c == '"') return true;
            if (static_cast<unsigned char>(c) < 0x20) return fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i_ >= s_.size()) return fail("unterminated escape");
            const char e = s_[i_++];
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
                    uint32_t cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xd800 && cp < 0xdc00) {  // surrogate pair
                        uint32_t lo;
                        if (i_ + 2 > s_.size() || s_[i_] != '\\' || s_[i_ + 1] != 'u')
                            return fail("unpaired surrogate");
                        i_ += 2;
                        if (!hex4(lo) || lo < 0xdc00 || lo >= 0xe000) return fail("unpaired surrogate");
                        cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
                    } else if (cp >= 0xdc00 && cp < 0xe000) {
                        return fail("unpaired surrogate");
                    }
                    utf8(out, cp);
                    break;
                }
                default: return fail("invalid escape");
            }
        }
    }
    bool number(Json &v) {
        const size_t start = i_;
        bool integer = true;
        if (s_[i_] == '-') ++i_;
        if (i_ >= s_.size()) return fail("invalid number");
        if (s_[i_] == '0') {
            ++i_;
        } else if (s_[i_] >= '1' && s_[i_] <= '9') {
            while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
        } else {
            return fail("invalid number");
        }
        if (i_ < s_.size() && s_[i_] == '.') {
            integer = false;
            ++i_;
            if (i_ >= s_.size() || s_[i_] < '0' || s_[i_] > '9') return fail("invalid number");
            while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
        }
        if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
            integer = false;
            ++i_;
            if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
            if (i_ >= s_.size() || s_[i_] < '0' || s_[i_] > '9') return fail("invalid number");
            while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
        }
        const std::string text = s_.substr(start, i_ - start);
        v.kind = Json::Number;
        if (integer) {
            // Integer literals keep their integer value (so -0 is +0), as a
            // 64-bit integer where it fits.
            errno = 0;
            if (text[0] == '-') {
                const long long x = std::strtoll(text.c_str(), nullptr, 10);
                if (errno == 0) {
                    v.num = double(x);
                    return true;
                }
            } else {
                const unsigned long long x = std::strtoull(text.c_str(), nullptr, 10);
                if (errno == 0) {
                    v.num = double(x);
                    return true;
                }
            }
        }
        v.num = std::strtod(text.c_str(), nullptr);
        if (!std::isfinite(v.num)) return fail("number out of range");
        return true;
    }
};

bool read_file(const std::string &path, std::string &out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::ostringstream ss;
    ss << in.rdbuf();
    out = ss.str();
    return true;
}

bool file_exists(const std::string &path) {
    struct stat st;
    return ::stat(path.c_str(), &st) == 0;
}

std::string parent_of(const std::string &path) {
    const size_t k = path.find_last_of('/');
    if (k == std::string::npos) return "";
    if (k == 0) return "/";
    return path.substr(0, k);
}

std::string join(const std::string &dir, const std::string &rel) {
    if (dir.empty()) return rel;
    if (dir.back() == '/') return dir + rel;
    return dir + "/" + rel;
}

std::string file_name(const std::string &path) {
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? path : path.substr(k + 1);
}

// Path::file_stem / extension: the name without its last ".ext" (a leading
// dot alone is part of the stem).
std::string stem_of(const std::string &path) {
    const std::string name = file_name(path);
    const size_t dot = name.find_last_of('.');
    return (dot == std::string::npos || dot == 0) ? name : name.substr(0, dot);
}
std::string extension_of(const std::string &path) {
    const std::string name = file_name(path);
    const size_t dot = name.find_last_of('.');
    return (dot == std::string::npos || dot == 0) ? "" : name.substr(dot + 1);
}

// ---------------------------------------------------------------------------
// Value conversions (src/scene.rs:780-808).
// ---------------------------------------------------------------------------
float number_f32(const Json *v, float fallback) {
    const double *d = v ? v->number() : nullptr;
    return d ? float(*d) : fallback;
}

// f64 -> u32 as the reference converts: truncation, saturating, NaN to 0.
uint32_t saturate_u32(double d) {
    if (!(d > 0.0)) return 0u;
    if (d >= 4294967295.0) return 0xffffffffu;
    return uint32_t(d);
}
uint32_t number_u32(const Json *v, uint32_t fallback) {
    const double *d = v ? v->number() : nullptr;
    return d ? saturate_u32(*d) : fallback;
}

void copy_float_array(float *target, size_t n, const Json &value) {
    if (!value.is_array()) return;
    for (size_t k = 0; k < n && k < value.items.size(); ++k) target[k] = number_f32(&value.items[k], target[k]);
}

void apply_float_object(float *target, const Json &value, const char *const *keys, size_t n) {
    if (!value.is_object()) return;
    for (size_t k = 0; k < n; ++k) target[k] = number_f32(value.get(keys[k]), target[k]);
}

bool bool_or_false(const Json *v) { return v && v->kind == Json::Bool && v->b; }

// ---------------------------------------------------------------------------
// Config pieces.
// ---------------------------------------------------------------------------
void apply_camera(FptRenderConfig &c, const Json &v) {  // scene.rs:810-824
    if (!v.is_object()) return;
    if (const Json *p = v.get("position")) copy_float_array(c.camera_position, 3, *p);
    if (const Json *p = v.get("yaw_pitch")) copy_float_array(c.camera_yaw_pitch, 2, *p);
    c.camera_roll = number_f32(v.get("roll"), c.camera_roll);
    c.camera_fov = number_f32(v.get("fov"), c.camera_fov);
    c.camera_dof = number_f32(v.get("dof"), c.camera_dof);
    c.focus_distance = number_f32(v.get("focus_distance"), c.focus_distance);
}

bool sdf_id_from_preset(const std::string &path, uint32_t &id) {  // scene.rs:832-848
    static const std::pair<const char *, uint32_t> table[] = {
        {"Cornell_Box", FPT_SDF_CORNELL_BOX},        {"Glass_Ball", FPT_SDF_GLASS_BALL},
        {"Ball_Fractal", FPT_SDF_BALL_FRACTAL},      {"Cage_Fractal", FPT_SDF_CAGE_FRACTAL},
        {"IFS_Fractal", FPT_SDF_IFS_FRACTAL},        {"Mandelbox_Fractal", FPT_SDF_MANDELBOX_FRACTAL},
        {"Menger_Sponge", FPT_SDF_MENGER_SPONGE},    {"Tower_Fractal", FPT_SDF_TOWER_FRACTAL},
        {"Tree_Fractal", FPT_SDF_TREE_FRACTAL},      {"README_Cornell", FPT_SDF_README_CORNELL},
        {"README_Glass", FPT_SDF_README_GLASS},
    };
    const std::string s = stem_of(path);
    for (const auto &e : table)
        if (s == e.first) {
            id = e.second;
            return true;
        }
    if (extension_of(path) == "fpt") {
        id = FPT_SDF_PROGRAM;
        return true;
    }
    return false;
}

std::string resolve_preset_path(const std::string &scene, const std::string &root, const std::string &preset) {
    if (!preset.empty() && preset[0] == '/') return preset;  // scene.rs:850-862
    const std::string parent = parent_of(scene);
    const std::string local = join(parent, preset);
    if (file_exists(local)) return local;
    return join(root, preset);
}

std::string resolve_scene_path(const std::string &scene, const std::string &value) {  // scene.rs:864-871
    if (!value.empty() && value[0] == '/') return value;
    const std::string parent = parent_of(scene);
    return join(parent.empty() ? "." : parent, value);
}

void set_hdri_path(FptRenderConfig &c, const std::string &path) {  // scene.rs:873-879
    std::memset(c.hdri_path, 0, sizeof(c.hdri_path));
    const size_t n = std::min(path.size(), sizeof(c.hdri_path) - 1);
    std::memcpy(c.hdri_path, path.data(), n);
    c.hdri_enabled = n > 0 ? 1u : 0u;
}

bool program_opcode(const std::string &name, uint32_t &op) {  // scene.rs:881-897
    static const std::pair<const char *, uint32_t> table[] = {
        {"abs", FPT_SDF_OP_ABS},           {"translate", FPT_SDF_OP_TRANSLATE}, {"scale", FPT_SDF_OP_SCALE},
        {"rotate_x", FPT_SDF_OP_ROTATE_X}, {"rotate_y", FPT_SDF_OP_ROTATE_Y},   {"rotate_z", FPT_SDF_OP_ROTATE_Z},
        {"repeat", FPT_SDF_OP_REPEAT},     {"sort_desc", FPT_SDF_OP_SORT_DESC}, {"sphere", FPT_SDF_OP_SPHERE},
        {"box", FPT_SDF_OP_BOX},           {"plane", FPT_SDF_OP_PLANE},         {"orbit_add", FPT_SDF_OP_ORBIT_ADD},
    };
    for (const auto &e : table)
        if (name == e.first) {
            op = e.second;
            return true;
        }
    return false;
}

uint32_t combine_mode(const Json *v) {  // scene.rs:899-905
    const char *s = v ? v->string() : nullptr;
    if (s && std::strcmp(s, "intersection") == 0) return 1u;
    if (s && std::strcmp(s, "subtract") == 0) return 2u;
    return 0u;
}

bool apply_gradient(FptRenderConfig &c, const Json &values, char *error, size_t error_len) {  // scene.rs:907-925
    if (values.items.size() > FPT_SDF_GRADIENT_MAX_STOPS) {
        seterr(error, error_len, "too many gradient stops");
        return false;
    }
    c.gradient_count = uint32_t(values.items.size());
    const float divisor = float(std::max<size_t>(values.items.size() > 0 ? values.items.size() - 1 : 0, 1));
    for (size_t k = 0; k < values.items.size(); ++k) {
        const Json &stop = values.items[k];
        if (!stop.is_object()) {
            seterr(error, error_len, "gradient stop must be an object");
            return false;
        }
        c.gradient_stops[k][0] = number_f32(stop.get("position"), float(k) / divisor);
        if (const Json *col = stop.get("color")) copy_float_array(&c.gradient_stops[k][1], 3, *col);
    }
    return true;
}

bool apply_fractal_style(FptRenderConfig &c, const Json &v, char *error, size_t error_len) {  // scene.rs:927-961
    if (!v.is_object()) {
        seterr(error, error_len, "style must be an object");
        return false;
    }
    if (const char *mode = v.get("mode") ? v.get("mode")->string() : nullptr) {
        if (std::strcmp(mode, "hsv") == 0) c.fractal_style_mode = 1u;
        else if (std::strcmp(mode, "solid") == 0) c.fractal_style_mode = 2u;
        else if (std::strcmp(mode, "gradient") == 0) c.fractal_style_mode = 3u;
        else {
            seterr(error, error_len, "invalid style mode: %s", mode);
            return false;
        }
    }
    static const char *const keys[] = {"hue_offset", "hue_scale", "saturation", "value",
                                       "roughness",  "specular",  "emission",   "solid_mix"};
    for (size_t k = 0; k < 8; ++k) c.fractal_style[k] = number_f32(v.get(keys[k]), c.fractal_style[k]);
    if (const Json *col = v.get("solid_color")) copy_float_array(&c.fractal_style[8], 3, *col);
    if (const Json *g = v.get("gradient"))
        if (g->is_array() && !apply_gradient(c, *g, error, error_len)) return false;
    return true;
}

bool apply_sdf_program(FptRenderConfig &c, const Json &v, char *error, size_t error_len) {  // scene.rs:963-1021
    if (!v.is_object()) {
        seterr(error, error_len, "sdf_program must be an object");
        return false;
    }
    const Json *ops = v.get("operations");
    if (!ops || !ops->is_array()) {
        seterr(error, error_len, "sdf_program.operations is required");
        return false;
    }
    if (ops->items.empty() || ops->items.size() > FPT_SDF_PROGRAM_MAX_OPS) {
        seterr(error, error_len, "invalid SDF operation count");
        return false;
    }
    c.sdf_id = FPT_SDF_PROGRAM;
    c.sdf_program_count = uint32_t(ops->items.size());
    for (size_t k = 0; k < ops->items.size(); ++k) {
        const Json &o = ops->items[k];
        if (!o.is_object()) {
            seterr(error, error_len, "SDF operation must be an object");
            return false;
        }
        const char *name = o.get("op") ? o.get("op")->string() : nullptr;
        if (!name) {
            seterr(error, error_len, "SDF operation is missing op");
            return false;
        }
        FptSdfInstruction ins{};
        if (!program_opcode(name, ins.opcode)) {
            seterr(error, error_len, "unsupported SDF operation: %s", name);
            return false;
        }
        ins.flags = combine_mode(o.get("combine"));
        const Json *data = o.get("data");
        if (!data) data = o.get("value");
        if (data) copy_float_array(ins.data, 4, *data);
        ins.data[0] = number_f32(o.get("radius"), ins.data[0]);
        ins.data[0] = number_f32(o.get("angle"), ins.data[0]);
        ins.data[0] = number_f32(o.get("factor"), ins.data[0]);
        ins.data[3] = number_f32(o.get("orbit_weight"), ins.data[3]);
        c.sdf_program[k] = ins;
    }
    if (const Json *m = v.get("material")) {
        if (m->is_object()) {
            if (const Json *col = m->get("color")) copy_float_array(c.program_material, 3, *col);
            static const char *const keys[] = {"roughness", "specular", "translucency", "ior", "emission"};
            for (size_t k = 0; k < 5; ++k) c.program_material[k + 3] = number_f32(m->get(keys[k]), c.program_material[k + 3]);
            const char *mode = m->get("mode") ? m->get("mode")->string() : nullptr;
            if (mode && std::strcmp(mode, "hsv_orbit") == 0) c.material_mode = 1u;
            else if (mode && std::strcmp(mode, "gradient") == 0) c.material_mode = 2u;
        }
    }
    if (const Json *g = v.get("gradient"))
        if (g->is_array() && !apply_gradient(c, *g, error, error_len)) return false;
    return true;
}

// ---------------------------------------------------------------------------
// Typed-program passes (scene.rs:1023-1530).
// ---------------------------------------------------------------------------
uint32_t fbits(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    return b;
}

bool same_metadata(const FptSdfInstruction &a, const FptSdfInstruction &b) {
    return a.flags == b.flags && a.material_index == b.material_index && a._pad0 == b._pad0;
}

bool same_bits(const float *a, const float *b, size_t n) {
    for (size_t k = 0; k < n; ++k)
        if (fbits(a[k]) != fbits(b[k])) return false;
    return true;
}

bool is_noop(const FptSdfInstruction &i) {
    switch (i.opcode) {
        case FPT_SDF_OP_TRANSLATE: return i.data[0] == 0.0f && i.data[1] == 0.0f && i.data[2] == 0.0f;
        case FPT_SDF_OP_SCALE: return i.data[0] == 1.0f || std::fabs(i.data[0]) <= 1.0e-6f;
        case FPT_SDF_OP_ROTATE_X:
        case FPT_SDF_OP_ROTATE_Y:
        case FPT_SDF_OP_ROTATE_Z: return i.data[0] == 0.0f;
        case FPT_SDF_OP_ORBIT_ADD:
            return i.data[3] == 0.0f || (i.data[0] == 0.0f && i.data[1] == 0.0f && i.data[2] == 0.0f);
        default: return false;
    }
}

// Merge `ins` into the preceding transform of the same kind.
bool fold_transform(FptSdfInstruction &prev, const FptSdfInstruction &ins) {
    if (prev.opcode != ins.opcode || !same_metadata(prev, ins)) return false;
    switch (ins.opcode) {
        case FPT_SDF_OP_TRANSLATE: {
            if (fbits(prev.data[3]) != fbits(ins.data[3])) return false;
            const float f[3] = {prev.data[0] + ins.data[0], prev.data[1] + ins.data[1], prev.data[2] + ins.data[2]};
            if (!std::isfinite(f[0]) || !std::isfinite(f[1]) || !std::isfinite(f[2])) return false;
            prev.data[0] = f[0];
            prev.data[1] = f[1];
            prev.data[2] = f[2];
            return true;
        }
        case FPT_SDF_OP_SCALE: {
            if (!same_bits(prev.data + 1, ins.data + 1, 3)) return false;
            const float f = prev.data[0] * ins.data[0];
            if (!std::isfinite(f) || !(std::fabs(f) > 1.0e-6f)) return false;
            prev.data[0] = f;
            return true;
        }
        case FPT_SDF_OP_ROTATE_X:
        case FPT_SDF_OP_ROTATE_Y:
        case FPT_SDF_OP_ROTATE_Z: {
            if (!same_bits(prev.data + 1, ins.data + 1, 3)) return false;
            const float f = prev.data[0] + ins.data[0];
            if (!std::isfinite(f)) return false;
            prev.data[0] = f;
            return true;
        }
        default: return false;
    }
}

// An identical hard-union or intersection primitive right after its twin
// (spheres and boxes only without an orbit weight).
bool redundant_primitive(const FptSdfInstruction &prev, const FptSdfInstruction &ins) {
    const bool prim = ins.opcode == FPT_SDF_OP_SPHERE || ins.opcode == FPT_SDF_OP_BOX || ins.opcode == FPT_SDF_OP_PLANE;
    if (prev.opcode != ins.opcode || !prim || !same_metadata(prev, ins) || !same_bits(prev.data, ins.data, 4) ||
        !(ins.flags == 0u || ins.flags == 1u))
        return false;
    return ins.opcode == FPT_SDF_OP_PLANE || ins.data[3] == 0.0f;
}

uint32_t program_len(const FptRenderConfig &c) {
    return std::min<uint32_t>(c.sdf_program_count, FPT_SDF_PROGRAM_MAX_OPS);
}

void rotate_rows(float rows[3][4], int a, int b, float angle) {  // scene.rs:1186-1194
    const float s = std::sin(angle), co = std::cos(angle);
    float ra[4], rb[4];
    std::memcpy(ra, rows[a], sizeof(ra));
    std::memcpy(rb, rows[b], sizeof(rb));
    for (int k = 0; k < 4; ++k) {
        rows[a][k] = co * ra[k] - s * rb[k];
        rows[b][k] = s * ra[k] + co * rb[k];
    }
}

bool finite_rows(const float rows[3][4]) {
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 4; ++k)
            if (!std::isfinite(rows[r][k])) return false;
    return true;
}

// The affine part of a program prefix: transform steps update `rows` (world
// point -> primitive frame) and the distance scale; false for an opcode that
// is not affine.
bool affine_step(const FptSdfInstruction &ins, float rows[3][4], float &scale_acc) {
    switch (ins.opcode) {
        case FPT_SDF_OP_TRANSLATE:
            for (int r = 0; r < 3; ++r) rows[r][3] -= ins.data[r];
            return true;
        case FPT_SDF_OP_SCALE: {
            const float s = std::fabs(ins.data[0]) > 1.0e-6f ? ins.data[0] : 1.0f;
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 4; ++k) rows[r][k] *= s;
            scale_acc *= std::fabs(s);
            return true;
        }
        case FPT_SDF_OP_ROTATE_X: rotate_rows(rows, 1, 2, ins.data[0]); return true;
        case FPT_SDF_OP_ROTATE_Y: rotate_rows(rows, 0, 2, ins.data[0]); return true;
        case FPT_SDF_OP_ROTATE_Z: rotate_rows(rows, 0, 1, ins.data[0]); return true;
        default: return false;
    }
}

// Plane data with a unit normal (offset unchanged); false for a degenerate or
// non-finite normal.
bool normalised_plane(float data[4]) {
    const float len = std::sqrt(data[0] * data[0] + data[1] * data[1] + data[2] * data[2]);
    if (!std::isfinite(len) || len <= 1.0e-8f) return false;
    data[0] /= len;
    data[1] /= len;
    data[2] /= len;
    return true;
}

FptPrimitiveInstance instance_of(const FptSdfInstruction &ins, const float data[4], const float rows[3][4],
                                 float scale_acc, uint32_t source, uint32_t mode) {
    FptPrimitiveInstance p{};
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 4; ++k) p.transform[r * 4 + k] = rows[r][k];
    std::memcpy(p.data, data, sizeof(p.data));
    p.opcode = ins.opcode;
    p.distance_scale = scale_acc;
    p.source_instruction = source;
    p._pad0 = mode;
    return p;
}

bool same_instance(const FptPrimitiveInstance &a, const FptPrimitiveInstance &b) {  // scene.rs:1196-1209
    return a.opcode == b.opcode && fbits(a.distance_scale) == fbits(b.distance_scale) &&
           same_bits(a.transform, b.transform, 12) && same_bits(a.data, b.data, 4);
}

bool finite4(const float *d) {
    return std::isfinite(d[0]) && std::isfinite(d[1]) && std::isfinite(d[2]) && std::isfinite(d[3]);
}

}  // namespace

void scene_default_config(FptRenderConfig &c) {  // scene.rs:743-778
    std::memset(&c, 0, sizeof(c));
    c.width = 768;
    c.height = 768;
    c.samples = 64;
    c.sdf_id = FPT_SDF_CORNELL_BOX;
    c.sdf_rr_start = 3.0f;
    c.sdf_rr_min_prob = 0.2f;
    c.mandel_iteration_scale = 1.0f;
    c.sdf_chunk_samples = 8;
    const float cam[3] = {0.1f, 0.1f, -5.0f};
    std::memcpy(c.camera_position, cam, sizeof(cam));
    c.camera_fov = 90.0f;
    c.camera_dof = 0.01f;
    const float render[8] = {5.0f, 312.0f, 0.0005f, 0.0005f, 1000.0f, 0.25f, 0.0f, 0.0f};
    std::memcpy(c.render, render, sizeof(render));
    const float world[7] = {0.0f, 1.0f, 120.0f, 30.0f, 1.0f, 1.0f, 0.0f};
    std::memcpy(c.world, world, sizeof(world));
    for (float &v : c.world_one_color) v = 1.0f;
    const float sun[5] = {0.0f, 120.0f, 30.0f, 1.0f, 0.0f};
    std::memcpy(c.sun, sun, sizeof(sun));
    for (float &v : c.sun_color) v = 1.0f;
    for (float &v : c.background_gradient) v = 1.0f;
    const float post[7] = {0.0f, 1.0f, 0.0f, 1.0f, 1.0f, 0.0f, 0.0f};
    std::memcpy(c.post, post, sizeof(post));
    const float mat[8] = {1.0f, 1.0f, 1.0f, 1.0f, 0.0f, 0.0f, 1.5f, 0.0f};
    std::memcpy(c.program_material, mat, sizeof(mat));
    const float style[12] = {0.0f, 1.0f, 0.5f, 1.0f, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f, 0.0f};
    std::memcpy(c.fractal_style, style, sizeof(style));
    c.voxel_resolution = 128;
    c.voxel_storage = FPT_VOXEL_STORAGE_SPARSE_BRICKS;
    for (float &v : c.voxel_bounds_min) v = -4.0f;
    for (float &v : c.voxel_bounds_max) v = 4.0f;
    c.voxel_surface_band = 1.0f;
    c.bound_grid_resolution = 32;
    c.bound_grid_profile_stride = 4;
    c.focus_distance = 0.0f;
}

void scene_optimize_program(FptRenderConfig &c) {  // scene.rs:1124-1157
    std::vector<FptSdfInstruction> out;
    const uint32_t n = program_len(c);
    for (uint32_t k = 0; k < n; ++k) {
        const FptSdfInstruction ins = c.sdf_program[k];
        if (is_noop(ins)) continue;
        if (!out.empty()) {
            FptSdfInstruction &prev = out.back();
            if ((ins.opcode == FPT_SDF_OP_ABS || ins.opcode == FPT_SDF_OP_SORT_DESC) && prev.opcode == ins.opcode) continue;
            if (redundant_primitive(prev, ins)) continue;
            if (fold_transform(prev, ins)) {
                if (is_noop(prev)) out.pop_back();
                continue;
            }
        }
        out.push_back(ins);
    }
    std::memset(c.sdf_program, 0, sizeof(c.sdf_program));
    std::copy(out.begin(), out.end(), c.sdf_program);
    c.sdf_program_count = uint32_t(out.size());
}

void scene_split_program(FptRenderConfig &c, bool enabled) {  // scene.rs:1159-1184
    std::memset(c.sdf_shading_program, 0, sizeof(c.sdf_shading_program));
    const uint32_t n = program_len(c);
    std::copy(c.sdf_program, c.sdf_program + n, c.sdf_shading_program);
    c.sdf_shading_program_count = n;
    c.sdf_geometry_split = enabled ? 1u : 0u;
    if (!enabled) return;
    std::vector<FptSdfInstruction> geom;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t op = c.sdf_program[k].opcode;
        if (op == FPT_SDF_OP_ORBIT_ADD || op == FPT_SDF_OP_UNION || op == FPT_SDF_OP_INTERSECTION ||
            op == FPT_SDF_OP_SUBTRACT)
            continue;
        geom.push_back(c.sdf_program[k]);
    }
    std::memset(c.sdf_program, 0, sizeof(c.sdf_program));
    std::copy(geom.begin(), geom.end(), c.sdf_program);
    c.sdf_program_count = uint32_t(geom.size());
}

bool scene_compile_canonical(FptRenderConfig &c) {  // scene.rs:1211-1358
    c.sdf_canonical_count = 0;
    c.sdf_canonical_source_count = 0;
    c.sdf_canonical_transform_count = 0;
    std::memset(c.sdf_canonical_primitives, 0, sizeof(c.sdf_canonical_primitives));
    std::memset(c.sdf_canonical_transforms, 0, sizeof(c.sdf_canonical_transforms));
    std::memset(c.sdf_indexed_primitives, 0, sizeof(c.sdf_indexed_primitives));
    if (c.sdf_id != FPT_SDF_PROGRAM || c.sdf_program_count == 0) return false;
    float rows[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
    float scale_acc = 1.0f;
    std::vector<FptPrimitiveInstance> prims;
    const uint32_t n = program_len(c);
    for (uint32_t k = 0; k < n; ++k) {
        const FptSdfInstruction &ins = c.sdf_program[k];
        if (!finite4(ins.data)) return false;
        const uint32_t op = ins.opcode;
        if (op == FPT_SDF_OP_SPHERE || op == FPT_SDF_OP_BOX || op == FPT_SDF_OP_PLANE) {
            if (ins.flags > 2u || prims.size() >= FPT_SDF_FLAT_UNION_MAX_PRIMITIVES) return false;
            c.sdf_canonical_source_count += 1;
            float data[4];
            std::memcpy(data, ins.data, sizeof(data));
            if (op == FPT_SDF_OP_PLANE) {
                const float len = std::sqrt(data[0] * data[0] + data[1] * data[1] + data[2] * data[2]);
                if (len <= 1.0e-8f) return false;
                data[0] /= len;
                data[1] /= len;
                data[2] /= len;
            }
            // The first primitive's combine flag is ignored (it starts the
            // fold); later ones keep theirs.
            const FptPrimitiveInstance p = instance_of(ins, data, rows, scale_acc, k, prims.empty() ? 0u : ins.flags);
            // A repeated primitive under an all-union (all-intersection)
            // program adds nothing: min (max) is idempotent.
            bool dup = false;
            if (ins.flags == 0u || ins.flags == 1u) {
                bool uniform = true;
                for (size_t q = 1; q < prims.size(); ++q) uniform = uniform && prims[q]._pad0 == ins.flags;
                bool any = false;
                for (const auto &q : prims) any = any || same_instance(q, p);
                dup = uniform && any;
            }
            if (!dup) prims.push_back(p);
        } else if (op == FPT_SDF_OP_UNION || op == FPT_SDF_OP_INTERSECTION || op == FPT_SDF_OP_SUBTRACT ||
                   op == FPT_SDF_OP_ORBIT_ADD) {
        } else if (!affine_step(ins, rows, scale_acc)) {
            return false;
        }
        if (!std::isfinite(scale_acc) || !finite_rows(rows)) return false;
    }
    if (prims.empty()) return false;
    std::copy(prims.begin(), prims.end(), c.sdf_canonical_primitives);
    c.sdf_canonical_count = uint32_t(prims.size());
    // Shared transforms (bit-identical rows and scale) are stored once.
    std::vector<FptAffineTransform> xf;
    for (size_t k = 0; k < prims.size(); ++k) {
        const FptPrimitiveInstance &p = prims[k];
        size_t idx = xf.size();
        for (size_t q = 0; q < xf.size(); ++q)
            if (fbits(xf[q].distance_scale) == fbits(p.distance_scale) && same_bits(xf[q].transform, p.transform, 12)) {
                idx = q;
                break;
            }
        if (idx == xf.size()) {
            FptAffineTransform t{};
            std::memcpy(t.transform, p.transform, sizeof(t.transform));
            t.distance_scale = p.distance_scale;
            xf.push_back(t);
        }
        FptIndexedPrimitive ip{};
        std::memcpy(ip.data, p.data, sizeof(ip.data));
        ip.opcode = p.opcode;
        ip.transform_index = uint32_t(idx);
        ip.source_instruction = p.source_instruction;
        ip.combine_mode = p._pad0;
        c.sdf_indexed_primitives[k] = ip;
    }
    c.sdf_canonical_transform_count = uint32_t(xf.size());
    std::copy(xf.begin(), xf.end(), c.sdf_canonical_transforms);
    return true;
}

bool scene_lower_flat_union(FptRenderConfig &c) {  // scene.rs:1360-1446
    c.sdf_flat_union_count = 0;
    std::memset(c.sdf_flat_union_instances, 0, sizeof(c.sdf_flat_union_instances));
    if (c.sdf_id != FPT_SDF_PROGRAM || c.sdf_program_count == 0) return false;
    float rows[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
    float scale_acc = 1.0f;
    std::vector<FptPrimitiveInstance> inst;
    const uint32_t n = program_len(c);
    for (uint32_t k = 0; k < n; ++k) {
        const FptSdfInstruction &ins = c.sdf_program[k];
        if (!finite4(ins.data)) return false;
        const uint32_t op = ins.opcode;
        if (op == FPT_SDF_OP_SPHERE || op == FPT_SDF_OP_BOX || op == FPT_SDF_OP_PLANE) {
            if (ins.flags != 0u || inst.size() >= FPT_SDF_FLAT_UNION_MAX_PRIMITIVES) return false;
            float data[4];
            std::memcpy(data, ins.data, sizeof(data));
            if (op == FPT_SDF_OP_PLANE && !normalised_plane(data)) return false;
            inst.push_back(instance_of(ins, data, rows, scale_acc, k, 0u));
        } else if (op == FPT_SDF_OP_ORBIT_ADD || op == FPT_SDF_OP_UNION) {
        } else if (!affine_step(ins, rows, scale_acc)) {
            return false;
        }
        if (!std::isfinite(scale_acc) || !finite_rows(rows)) return false;
    }
    if (inst.empty()) return false;
    std::copy(inst.begin(), inst.end(), c.sdf_flat_union_instances);
    c.sdf_flat_union_count = uint32_t(inst.size());
    return true;
}

bool scene_lower_typed_soa(FptRenderConfig &c) {  // scene.rs:1448-1530
    std::memset(&c.sdf_typed_soa, 0, sizeof(c.sdf_typed_soa));
    if (c.sdf_id != FPT_SDF_PROGRAM || c.sdf_program_count == 0) return false;
    FptTypedSoAProgram &t = c.sdf_typed_soa;
    float tr[3] = {0.f, 0.f, 0.f};
    size_t count = 0;
    const uint32_t n = program_len(c);
    for (uint32_t k = 0; k < n; ++k) {
        const FptSdfInstruction &ins = c.sdf_program[k];
        if (!finite4(ins.data)) return false;
        switch (ins.opcode) {
            case FPT_SDF_OP_TRANSLATE:
                for (int a = 0; a < 3; ++a) tr[a] += ins.data[a];
                break;
            case FPT_SDF_OP_SPHERE: {
                if (ins.flags != 0u || count >= FPT_SDF_FLAT_UNION_MAX_PRIMITIVES) return false;
                const uint32_t i = t.sphere_count++;
                t.sphere_x[i] = tr[0];
                t.sphere_y[i] = tr[1];
                t.sphere_z[i] = tr[2];
                t.sphere_radius[i] = ins.data[0];
                t.sphere_source[i] = k;
                ++count;
                break;
            }
            case FPT_SDF_OP_BOX: {
                if (ins.flags != 0u || count >= FPT_SDF_FLAT_UNION_MAX_PRIMITIVES) return false;
                const uint32_t i = t.box_count++;
                t.box_x[i] = tr[0];
                t.box_y[i] = tr[1];
                t.box_z[i] = tr[2];
                t.box_half_x[i] = std::fabs(ins.data[0]);
                t.box_half_y[i] = std::fabs(ins.data[1]);
                t.box_half_z[i] = std::fabs(ins.data[2]);
                t.box_source[i] = k;
                ++count;
                break;
            }
            case FPT_SDF_OP_PLANE: {
                if (ins.flags != 0u || count >= FPT_SDF_FLAT_UNION_MAX_PRIMITIVES) return false;
                const float len = std::sqrt(ins.data[0] * ins.data[0] + ins.data[1] * ins.data[1] + ins.data[2] * ins.data[2]);
                if (!std::isfinite(len) || len <= 1.0e-8f) return false;
                const uint32_t i = t.plane_count++;
                t.plane_x[i] = ins.data[0];  // the normal stays unnormalised here
                t.plane_y[i] = ins.data[1];
                t.plane_z[i] = ins.data[2];
                t.plane_center_x[i] = tr[0];
                t.plane_center_y[i] = tr[1];
                t.plane_center_z[i] = tr[2];
                t.plane_offset[i] = ins.data[3];
                t.plane_source[i] = k;
                ++count;
                break;
            }
            case FPT_SDF_OP_ORBIT_ADD:
            case FPT_SDF_OP_UNION: break;
            default: return false;
        }
        if (!std::isfinite(tr[0]) || !std::isfinite(tr[1]) || !std::isfinite(tr[2])) return false;
    }
    return count > 0;
}

namespace {

bool append_op(FptRenderConfig &c, uint32_t op, float a, float b, float d, float e, char *error, size_t error_len) {
    if (c.sdf_program_count >= FPT_SDF_PROGRAM_MAX_OPS) {  // scene.rs:1532-1545
        seterr(error, error_len, "SDF program is too large");
        return false;
    }
    FptSdfInstruction ins{};
    ins.opcode = op;
    ins.data[0] = a;
    ins.data[1] = b;
    ins.data[2] = d;
    ins.data[3] = e;
    c.sdf_program[c.sdf_program_count++] = ins;
    return true;
}

// The preset's "Gradient(Orbit_Trap)" example as a typed program
// (scene.rs:1558-1581): ten fold/rotate/scale/translate rounds and a box.
bool compile_gradient_example(FptRenderConfig &c, char *error, size_t error_len) {
    c.sdf_id = FPT_SDF_PROGRAM;
    c.material_mode = 2u;
    const float angle1 = -9.83f + c.set_values[2] / 2.0f;
    const float angle2 = -1.16f + c.set_values[3] / 2.0f;
    const float shift[3] = {-3.508f + c.set_values[4] * 7.0f, -3.593f + c.set_values[5] * 7.0f,
                            3.295f + c.set_values[6] * 7.0f};
    for (int r = 0; r < 10; ++r) {
        if (!append_op(c, FPT_SDF_OP_ABS, 0, 0, 0, 0, error, error_len) ||
            !append_op(c, FPT_SDF_OP_ROTATE_Z, -angle1, 0, 0, 0, error, error_len) ||
            !append_op(c, FPT_SDF_OP_SORT_DESC, 0, 0, 0, 0, error, error_len) ||
            !append_op(c, FPT_SDF_OP_ROTATE_X, -angle2, 0, 0, 0, error, error_len) ||
            !append_op(c, FPT_SDF_OP_SCALE, 1.9f, 0, 0, 0, error, error_len) ||
            !append_op(c, FPT_SDF_OP_TRANSLATE, -shift[0], -shift[1], -shift[2], 0, error, error_len))
            return false;
    }
    return append_op(c, FPT_SDF_OP_BOX, 6.0f, 6.0f, 6.0f, 1.0f, error, error_len);
}

std::string trim_ws(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

// A float literal as the reference's parser accepts it: optional sign,
// decimal digits with an optional point and exponent, or inf / infinity /
// nan in any case; nothing else on the line.
bool parse_f32(const std::string &s, float &out) {
    if (s.empty()) return false;
    size_t k = 0;
    if (s[k] == '+' || s[k] == '-') ++k;
    std::string rest = s.substr(k);
    for (char &ch : rest) ch = char(std::tolower(static_cast<unsigned char>(ch)));
    if (rest == "inf" || rest == "infinity" || rest == "nan") {
        const float v = rest == "nan" ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
        out = s[0] == '-' ? -v : v;
        return true;
    }
    bool digits = false, point = false;
    size_t i = k;
    while (i < s.size() && (std::isdigit(static_cast<unsigned char>(s[i])) || (s[i] == '.' && !point))) {
        if (s[i] == '.') point = true;
        else digits = true;
        ++i;
    }
    if (!digits) return false;
    if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
        if (i >= s.size() || !std::isdigit(static_cast<unsigned char>(s[i]))) return false;
        while (i < s.size() && std::isdigit(static_cast<unsigned char>(s[i]))) ++i;
    }
    if (i != s.size()) return false;
    out = std::strtof(s.c_str(), nullptr);
    return true;
}

bool parse_vec3(const std::string &line, float v[3]) {  // scene.rs:1547-1556
    size_t a = 0, b = line.size();
    auto strip = [](char ch) { return std::isspace(static_cast<unsigned char>(ch)) || ch == '[' || ch == ']'; };
    while (a < b && strip(line[a])) ++a;
    while (b > a && strip(line[b - 1])) --b;
    const std::string body = line.substr(a, b - a);
    size_t start = 0;
    for (int k = 0; k < 3; ++k) {
        if (start > body.size()) return false;
        size_t comma = body.find(',', start);
        const std::string part = trim_ws(body.substr(start, comma == std::string::npos ? std::string::npos : comma - start));
        if (!parse_f32(part, v[k])) return false;
        if (comma == std::string::npos) {
            if (k < 2) return false;
            start = body.size() + 1;
        } else {
            start = comma + 1;
        }
    }
    return true;
}

}  // namespace

bool scene_load_fpt_settings(const std::string &path, FptRenderConfig &c, char *error, size_t error_len) {
    std::string data;  // scene.rs:1583-1617
    if (!read_file(path, data)) {
        seterr(error, error_len, "failed to read preset %s: %s", path.c_str(), std::strerror(errno ? errno : ENOENT));
        return false;
    }
    const size_t marker = data.find("SDF Settings:");
    if (marker == std::string::npos) return true;
    int mode = 0;
    size_t set_index = 0, vec_index = 0;
    std::istringstream lines(data.substr(marker));
    std::string raw;
    while (std::getline(lines, raw)) {
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        const std::string line = trim_ws(raw);
        if (line == "{") {
            ++mode;
            continue;
        }
        if (line == "}") continue;
        if (mode == 1 && set_index < 40) {
            float v;
            if (parse_f32(line, v)) c.set_values[set_index++] = v;
        } else if (mode == 2 && vec_index < 40) {
            float v[3];
            if (parse_vec3(line, v)) {
                std::memcpy(&c.vset_values[vec_index * 3], v, sizeof(v));
                ++vec_index;
            }
        }
    }
    if (c.sdf_id == FPT_SDF_PROGRAM && data.find("Gradient(Orbit_Trap)") != std::string::npos)
        return compile_gradient_example(c, error, error_len);
    return true;
}

bool scene_apply_program_json(FptRenderConfig &c, const std::string &text, char *error, size_t error_len) {
    Json v;
    std::string err;
    if (!JsonParser(text).parse(v, err)) {
        seterr(error, error_len, "%s", err.c_str());
        return false;
    }
    return apply_sdf_program(c, v, error, error_len);
}

bool scene_load(const std::string &scene_path, const std::string &fpt_root, const FptSceneOptions &opt,
                FptRenderConfig &c, std::string &output_name, char *error, size_t error_len) {
    if (extension_of(scene_path) == "fract") return fract_load(scene_path, opt, c, output_name, error, error_len);
    std::string text;  // scene.rs:1619-1976
    if (!read_file(scene_path, text)) {
        seterr(error, error_len, "failed to read scene %s: %s", scene_path.c_str(), std::strerror(errno ? errno : ENOENT));
        return false;
    }
    Json root;
    std::string err;
    if (!JsonParser(text).parse(root, err)) {
        seterr(error, error_len, "%s: %s", scene_path.c_str(), err.c_str());
        return false;
    }
    if (!root.is_object()) {
        seterr(error, error_len, "scene root must be an object");
        return false;
    }
    const char *preset = root.get("preset") ? root.get("preset")->string() : nullptr;
    if (!preset) {
        seterr(error, error_len, "scene is missing preset");
        return false;
    }
    const std::string preset_path = resolve_preset_path(scene_path, fpt_root, preset);
    scene_default_config(c);
    if (!sdf_id_from_preset(preset_path, c.sdf_id)) {
        seterr(error, error_len, "unsupported preset: %s", preset_path.c_str());
        return false;
    }
    c.preview = opt.preview;
    c.sdf_profile = opt.sdf_profile;
    c.glass_mode = opt.glass_mode;
    c.width = opt.width ? opt.width : number_u32(root.get("width"), 768);
    c.height = opt.height ? opt.height : number_u32(root.get("height"), 768);
    c.samples = opt.samples ? opt.samples : number_u32(root.get("samples"), 64);
    if (c.samples < 1u || c.samples > 512u) {
        seterr(error, error_len, "scene samples must be 1..512");
        return false;
    }
    if (const Json *v = root.get("camera")) apply_camera(c, *v);
    if (const Json *v = root.get("render")) {
        static const char *const keys[] = {"bounces",      "marching_steps",    "normal_epsilon",   "min_distance",
                                           "max_distance", "adaptive_marching", "error_protection", "random_mode"};
        apply_float_object(c.render, *v, keys, 8);
    }
    if (const Json *v = root.get("world")) {
        static const char *const keys[] = {"mode", "light_size", "rotation", "elevation", "power", "contrast",
                                           "gradient_background"};
        apply_float_object(c.world, *v, keys, 7);
        if (v->is_object()) {
            if (const Json *col = v->get("one_color")) copy_float_array(c.world_one_color, 3, *col);
            const Json *bg = v->get("background_gradient");
            if (bg && bg->is_array() && bg->items.size() >= 2) {
                copy_float_array(c.background_gradient, 3, bg->items[0]);
                copy_float_array(c.background_gradient + 3, 3, bg->items[1]);
            }
        }
    }
    if (const Json *v = root.get("sun")) {
        static const char *const keys[] = {"enabled", "rotation", "elevation", "power", "softness"};
        apply_float_object(c.sun, *v, keys, 5);
        if (v->is_object())
            if (const Json *col = v->get("color")) copy_float_array(c.sun_color, 3, *col);
    }
    if (const Json *v = root.get("post")) {
        static const char *const keys[] = {"tone_map", "exposure", "brightness",  "saturation",
                                           "contrast", "chromatic_aberration", "highlights"};
        apply_float_object(c.post, *v, keys, 7);
    }
    if (const Json *v = root.get("style"))
        if (!apply_fractal_style(c, *v, error, error_len)) return false;
    if (const Json *vox = root.get("voxel")) {
        if (vox->is_object()) {
            if (const Json *r = vox->get("resolution")) c.voxel_resolution = number_u32(r, c.voxel_resolution);
            if (const Json *b = vox->get("bounds_min")) copy_float_array(c.voxel_bounds_min, 3, *b);
            if (const Json *b = vox->get("bounds_max")) copy_float_array(c.voxel_bounds_max, 3, *b);
            c.voxel_surface_band = number_f32(vox->get("surface_band"), c.voxel_surface_band);
            if (const char *s = vox->get("storage") ? vox->get("storage")->string() : nullptr) {
                if (!std::strcmp(s, "dense")) c.voxel_storage = FPT_VOXEL_STORAGE_DENSE;
                else if (!std::strcmp(s, "sparse-bricks")) c.voxel_storage = FPT_VOXEL_STORAGE_SPARSE_BRICKS;
                else if (!std::strcmp(s, "template-bricks")) c.voxel_storage = FPT_VOXEL_STORAGE_TEMPLATE_BRICKS;
                else if (!std::strcmp(s, "sparse-bricks8")) c.voxel_storage = FPT_VOXEL_STORAGE_SPARSE_BRICKS8;
                else {
                    seterr(error, error_len, "invalid voxel storage mode: %s", s);
                    return false;
                }
            }
            // 32..512 as the reference; the 8^3-brick layout of this library
            // also takes 1024.
            const uint32_t max_res = c.voxel_storage == FPT_VOXEL_STORAGE_SPARSE_BRICKS8 ? 1024u : 512u;
            if (c.voxel_resolution < 32u || c.voxel_resolution > max_res) {
                seterr(error, error_len, "voxel resolution must be 32..%u", max_res);
                return false;
            }
            c.voxel_fill_interior = bool_or_false(vox->get("fill_interior")) ? 1u : 0u;
            if (const char *s = vox->get("coverage") ? vox->get("coverage")->string() : nullptr) {
                if (!std::strcmp(s, "legacy")) c.voxel_coverage_mode = FPT_VOXEL_COVERAGE_LEGACY;
                else if (!std::strcmp(s, "lipschitz")) c.voxel_coverage_mode = FPT_VOXEL_COVERAGE_LIPSCHITZ;
                else if (!std::strcmp(s, "interval")) c.voxel_coverage_mode = FPT_VOXEL_COVERAGE_INTERVAL;
                else if (!std::strcmp(s, "shell")) c.voxel_coverage_mode = FPT_VOXEL_COVERAGE_SHELL;
                else {
                    seterr(error, error_len, "invalid voxel coverage mode: %s", s);
                    return false;
                }
            }
            if (const char *s = vox->get("build") ? vox->get("build")->string() : nullptr) {
                if (!std::strcmp(s, "staging")) c.voxel_build_mode = FPT_VOXEL_BUILD_STAGING;
                else if (!std::strcmp(s, "direct")) c.voxel_build_mode = FPT_VOXEL_BUILD_DIRECT;
                else {
                    seterr(error, error_len, "invalid voxel build mode: %s", s);
                    return false;
                }
            }
            c.voxel_brick_rejection = bool_or_false(vox->get("brick_rejection")) ? 1u : 0u;
            if (const char *s = vox->get("leaf_refinement") ? vox->get("leaf_refinement")->string() : nullptr) {
                if (!std::strcmp(s, "none")) c.voxel_leaf_refinement = FPT_VOXEL_LEAF_REFINEMENT_NONE;
                else if (!std::strcmp(s, "secant-bisection"))
                    c.voxel_leaf_refinement = FPT_VOXEL_LEAF_REFINEMENT_SECANT_BISECTION;
                else if (!std::strcmp(s, "restricted-trace"))
                    c.voxel_leaf_refinement = FPT_VOXEL_LEAF_REFINEMENT_RESTRICTED_TRACE;
                else if (!std::strcmp(s, "fixed-de")) c.voxel_leaf_refinement = FPT_VOXEL_LEAF_REFINEMENT_FIXED_DE;
                else {
                    seterr(error, error_len, "invalid voxel leaf refinement: %s", s);
                    return false;
                }
            }
            if (const char *s = vox->get("material") ? vox->get("material")->string() : nullptr) {
                if (!std::strcmp(s, "stored")) c.voxel_material_mode = FPT_VOXEL_MATERIAL_STORED;
                else if (!std::strcmp(s, "exact")) c.voxel_material_mode = FPT_VOXEL_MATERIAL_EXACT;
                else {
                    seterr(error, error_len, "invalid voxel material mode: %s", s);
                    return false;
                }
            }
            if (const char *s = vox->get("offset") ? vox->get("offset")->string() : nullptr) {
                if (!std::strcmp(s, "legacy")) c.voxel_offset_mode = FPT_VOXEL_OFFSET_LEGACY;
                else if (!std::strcmp(s, "precision")) c.voxel_offset_mode = FPT_VOXEL_OFFSET_PRECISION;
                else {
                    seterr(error, error_len, "invalid voxel offset mode: %s", s);
                    return false;
                }
            }
        }
    }
    for (int a = 0; a < 3; ++a)
        if (!(c.voxel_bounds_min[a] < c.voxel_bounds_max[a])) {
            seterr(error, error_len, "voxel bounds_max must be greater than bounds_min");
            return false;
        }
    if (!scene_load_fpt_settings(preset_path, c, error, error_len)) return false;
    if (const Json *v = root.get("sdf_program"))
        if (!apply_sdf_program(c, *v, error, error_len)) return false;
    c.sdf_program_source_count = c.sdf_program_count;
    c.sdf_program_optimization = opt.program_optimization;
    if (c.sdf_id == FPT_SDF_PROGRAM && opt.program_optimization == 1u) scene_optimize_program(c);
    if (c.sdf_id == FPT_SDF_PROGRAM) {
        scene_split_program(c, opt.geometry_split != 0u);
        if (opt.canonical_ir) scene_compile_canonical(c);
    }
    const bool sdf_backend = opt.renderer_backend == FPT_RENDERER_SDF;
    if (opt.topology_specialization && !sdf_backend) {
        seterr(error, error_len, "topology specialization currently requires the direct SDF renderer");
        return false;
    }
    if (opt.typed_soa == 1u && !sdf_backend) {
        seterr(error, error_len, "typed-SoA lowering currently requires the direct SDF renderer");
        return false;
    }
    if (opt.topology_specialization && opt.flat_union) {
        seterr(error, error_len, "topology specialization and flat-union lowering are separate evaluator modes");
        return false;
    }
    if (opt.topology_specialization && opt.typed_soa == 1u) {
        seterr(error, error_len, "topology specialization and typed-SoA lowering are separate evaluator modes");
        return false;
    }
    if (opt.flat_union && opt.typed_soa == 1u) {
        seterr(error, error_len, "flat-union and typed-SoA lowering are separate evaluator modes");
        return false;
    }
    if (opt.flat_union) scene_lower_flat_union(c);
    if (opt.typed_soa == 1u) {
        if (!scene_lower_typed_soa(c)) {
            seterr(error, error_len,
                   "typed-SoA lowering requires a translation-only hard union of spheres, boxes, or planes");
            return false;
        }
    } else if (opt.typed_soa == 2u) {
        scene_lower_typed_soa(c);
    }
    if (c.sdf_id == FPT_SDF_PROGRAM && c.sdf_program_count == 0) {
        seterr(error, error_len, "preset '%s' needs a typed sdf_program", preset_path.c_str());
        return false;
    }
    if (const Json *w = root.get("world"))
        if (w->is_object())
            if (const char *h = w->get("hdri") ? w->get("hdri")->string() : nullptr)
                set_hdri_path(c, resolve_scene_path(scene_path, h));
    const char *out = root.get("output") ? root.get("output")->string() : nullptr;
    output_name = out ? std::string(out) : stem_of(preset_path) + ".png";
    return true;
}

}  // namespace fpt

// ---------------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------------
extern "C" {

int fpt_scene_default_config(FptRenderConfig *config) {
    if (!config) return 1;
    fpt::scene_default_config(*config);
    return 0;
}

int fpt_scene_load(const char *scene_path, const char *fpt_root, const FptSceneOptions *options,
                   FptRenderConfig *config, char *output_name, size_t output_name_len, char *error, size_t error_len) {
    if (!scene_path || !config) {
        fpt::seterr(error, error_len, "missing scene path or config");
        return 1;
    }
    FptSceneOptions opt{};
    if (options) opt = *options;
    FptRenderConfig c;
    std::string name;
    try {  // no C++ exception (allocation failure) may cross the C ABI
        if (!fpt::scene_load(scene_path, fpt_root ? fpt_root : ".", opt, c, name, error, error_len)) return 1;
    } catch (const std::exception &e) {
        fpt::seterr(error, error_len, "load scene %s: %s", scene_path, e.what());
        return 1;
    }
    *config = c;
    if (output_name && output_name_len) std::snprintf(output_name, output_name_len, "%s", name.c_str());
    return 0;
}

int fpt_scene_load_fpt_settings(const char *preset_path, FptRenderConfig *config, char *error, size_t error_len) {
    if (!preset_path || !config) {
        fpt::seterr(error, error_len, "missing preset path or config");
        return 1;
    }
    return fpt::scene_load_fpt_settings(preset_path, *config, error, error_len) ? 0 : 1;
}

int fpt_scene_apply_program_json(FptRenderConfig *config, const char *json, char *error, size_t error_len) {
    if (!config || !json) {
        fpt::seterr(error, error_len, "missing config or JSON");
        return 1;
    }
    return fpt::scene_apply_program_json(*config, json, error, error_len) ? 0 : 1;
}

void fpt_scene_optimize_program(FptRenderConfig *config) {
    if (config) fpt::scene_optimize_program(*config);
}

void fpt_scene_split_program(FptRenderConfig *config, uint32_t enabled) {
    if (config) fpt::scene_split_program(*config, enabled != 0u);
}

int fpt_scene_compile_canonical(FptRenderConfig *config) {
    return config && fpt::scene_compile_canonical(*config) ? 1 : 0;
}

int fpt_scene_lower_flat_union(FptRenderConfig *config) {
    return config && fpt::scene_lower_flat_union(*config) ? 1 : 0;
}

int fpt_scene_lower_typed_soa(FptRenderConfig *config) {
    return config && fpt::scene_lower_typed_soa(*config) ? 1 : 0;
}

}  // extern "C"
