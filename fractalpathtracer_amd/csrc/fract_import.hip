// Mandelbulber 2.x settings (.fract) import for the native Kaleidoscopic IFS
// field (formula 10): the settings document, the scene parameters the field
// and the camera use, and their placement in the config
// (src/mandelbulber.rs:226-1041 parse, 1177-1344 apply_to_config;
// src/scene.rs:1978-2037 load_mandelbulber_scene_config).  Other formulas
// and hybrids need the external Mandelbulber2 formula source and are
// rejected.  Host code only.
//
// Document format: "# version M.m" comment; "[section]" headers;
// "key value;" statements in [main_parameters] and [fractal_1..9] (other
// sections are skipped); numbers use a decimal comma or point; booleans are
// true / false; vectors are three numbers separated by spaces.
#pragma clang fp contract(off)

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "fpt_internal.h"

namespace fpt {
namespace {

constexpr double kWorldScale = 1024.0;       // mandelbulber.rs:23
constexpr double kDefaultFov = 53.13;        // mandelbulber.rs:13
constexpr uint32_t kDefaultMaxIter = 250;    // mandelbulber.rs:14
constexpr double kDefaultBailout = 100.0;    // mandelbulber.rs:15
constexpr int kIfsVectors = 9;               // mandelbulber.rs:16
constexpr int kFormulaSlots = 9;             // mandelbulber.rs:84
constexpr double kPi = 3.141592653589793;
// set_values / vset_values slots (mandelbulber.rs:25-83).
enum : int {
    P_WORLD_SCALE = 0, P_MAX_ITER = 1, P_BAILOUT = 2, P_IFS_SCALE = 3, P_ABS_MASK = 4, P_ENABLED_MASK = 5,
    P_OFFSET = 6, P_ROTATION = 9, P_DIRECTION = 12, P_INITIAL_W = 3, P_ADD_C = 4, P_JULIA_MODE = 5, P_JULIA_C = 6,
    P_CONST_MUL = 9, P_FORMULA_ID = 39,
    V_BOX_FOLD = 101, V_BOX_LIMIT = 102, V_BOX_VALUE = 103, V_SPH_FOLD = 104, V_SPH_OUTER = 105, V_SPH_INNER = 106,
    V_PROJECTION = 107, V_ADV_QUALITY = 108, V_ABS_MIN = 109, V_ABS_MAX = 110, V_REL_MIN = 111, V_REL_MAX = 112,
    V_DE_FACTOR = 113, V_SMOOTH = 114, V_DYN_THRESH = 115, V_THRESH_SCALE = 116, V_CONST_THRESH = 117,
    V_MIN_THRESH = 118, V_MAX_THRESH = 119, V_POSITION = 120, V_ROTATION = 123, V_REPEAT = 126, V_DETAIL = 129,
    V_REFINEMENT = 131, V_SPATIAL = 132,
};
constexpr float kIterationThresholdMode = 2.0f;

struct Fail {
    std::string msg;
};
[[noreturn]] void fail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Fail{buf};
}

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

struct Document {
    std::string version;
    std::map<std::string, std::map<std::string, std::string>> sections;

    const std::string *value(const std::string &sec, const std::string &key) const {
        auto s = sections.find(sec);
        if (s == sections.end()) return nullptr;
        auto v = s->second.find(key);
        return v == s->second.end() ? nullptr : &v->second;
    }
};

// A decimal literal with ',' or '.' as the decimal mark (mandelbulber.rs:339-344).
double parse_number(const std::string &text) {
    std::string t = text;
    for (char &c : t)
        if (c == ',') c = '.';
    if (t.empty()) fail("invalid decimal number '%s'", text.c_str());
    size_t i = 0;
    if (t[i] == '+' || t[i] == '-') ++i;
    std::string rest = t.substr(i);
    for (char &c : rest) c = char(std::tolower(static_cast<unsigned char>(c)));
    if (rest == "inf" || rest == "infinity" || rest == "nan") {
        const double v = rest == "nan" ? std::nan("") : INFINITY;
        return t[0] == '-' ? -v : v;
    }
    bool digits = false, point = false;
    size_t k = i;
    while (k < t.size() && (std::isdigit(static_cast<unsigned char>(t[k])) || (t[k] == '.' && !point))) {
        if (t[k] == '.') point = true;
        else digits = true;
        ++k;
    }
    if (!digits) fail("invalid decimal number '%s'", text.c_str());
    if (k < t.size() && (t[k] == 'e' || t[k] == 'E')) {
        ++k;
        if (k < t.size() && (t[k] == '+' || t[k] == '-')) ++k;
        if (k >= t.size() || !std::isdigit(static_cast<unsigned char>(t[k]))) fail("invalid decimal number '%s'", text.c_str());
        while (k < t.size() && std::isdigit(static_cast<unsigned char>(t[k]))) ++k;
    }
    if (k != t.size()) fail("invalid decimal number '%s'", text.c_str());
    return std::strtod(t.c_str(), nullptr);
}

Document parse_document(const std::string &src) {  // mandelbulber.rs:232-293
    Document d;
    bool have_version = false;
    std::string section;
    bool in_section = false;
    std::istringstream lines(src);
    std::string raw;
    size_t number = 0;
    while (std::getline(lines, raw)) {
        ++number;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        const std::string line = trim(raw);
        if (line.empty()) continue;
        if (line[0] == '#') {
            const std::string c = trim(line.substr(1));
            if (c.compare(0, 8, "version ") == 0) {
                d.version = trim(c.substr(8));
                have_version = true;
            }
            continue;
        }
        if (line.front() == '[' && line.back() == ']') {
            const std::string name = trim(line.substr(1, line.size() - 2));
            if (name.empty()) fail("empty section at line %zu", number);
            section = name;
            in_section = true;
            d.sections[name];
            continue;
        }
        if (!in_section) fail("parameter outside a section at line %zu", number);
        bool kept = section == "main_parameters";
        if (!kept && section.compare(0, 8, "fractal_") == 0) {
            const std::string idx = section.substr(8);
            kept = !idx.empty() && idx.find_first_not_of("0123456789") == std::string::npos && idx.size() < 6 &&
                   std::atoi(idx.c_str()) <= 9;
        }
        if (!kept) continue;
        if (line.back() != ';') fail("missing ';' at line %zu", number);
        const std::string stmt = line.substr(0, line.size() - 1);
        size_t sp = 0;
        while (sp < stmt.size() && !std::isspace(static_cast<unsigned char>(stmt[sp]))) ++sp;
        if (sp == stmt.size()) fail("missing parameter value at line %zu", number);
        const std::string key = trim(stmt.substr(0, sp)), value = trim(stmt.substr(sp));
        if (key.empty()) fail("empty parameter name at line %zu", number);
        if (value.empty()) continue;
        if (!d.sections[section].emplace(key, value).second)
            fail("duplicate parameter %s.%s at line %zu", section.c_str(), key.c_str(), number);
    }
    if (!have_version) fail("missing Mandelbulber version comment");
    return d;
}

double number(const Document &d, const char *sec, const std::string &key, double fallback) {
    const std::string *v = d.value(sec, key);
    if (!v) return fallback;
    try {
        return parse_number(*v);
    } catch (const Fail &) {
        fail("invalid %s.%s", sec, key.c_str());
    }
}

uint32_t integer(const Document &d, const char *sec, const std::string &key, uint32_t fallback) {
    const double v = number(d, sec, key, double(fallback));
    if (!std::isfinite(v) || v < 0.0) fail("%s.%s must be non-negative", sec, key.c_str());
    if (v != std::floor(v)) fail("%s.%s must be an integer", sec, key.c_str());
    if (v > 4294967295.0) fail("%s.%s is too large", sec, key.c_str());
    return uint32_t(v);
}

bool boolean(const Document &d, const char *sec, const std::string &key, bool fallback) {
    const std::string *v = d.value(sec, key);
    if (!v) return fallback;
    if (*v == "true") return true;
    if (*v == "false") return false;
    fail("invalid %s.%s boolean: %s", sec, key.c_str(), v->c_str());
}

struct Vec3 {
    double x, y, z;
};

Vec3 vec3(const Document &d, const char *sec, const std::string &key, Vec3 fallback) {
    const std::string *v = d.value(sec, key);
    if (!v) return fallback;
    std::istringstream ss(*v);
    std::vector<double> out;
    std::string tok;
    try {
        while (ss >> tok) out.push_back(parse_number(tok));
    } catch (const Fail &) {
        fail("invalid %s.%s", sec, key.c_str());
    }
    if (out.size() != 3) fail("invalid %s.%s: expected three vector components", sec, key.c_str());
    return {out[0], out[1], out[2]};
}

double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double vlen(Vec3 a) { return std::sqrt(dot(a, a)); }
Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 mul(Vec3 a, double k) { return {a.x * k, a.y * k, a.z * k}; }
Vec3 normalized(Vec3 v) {  // mandelbulber.rs:2505-2512
    const double m = vlen(v);
    if (!(m > 0.0) || !std::isfinite(m)) fail("vector must be finite and non-zero");
    return mul(v, 1.0 / m);
}
double to_radians(double deg) { return deg * (kPi / 180.0); }
double to_degrees(double rad) { return rad * (180.0 / kPi); }

bool version_before(const std::string &v, uint32_t major, uint32_t minor) {  // mandelbulber.rs:346-359
    const size_t dot1 = v.find('.');
    if (dot1 == std::string::npos) fail("invalid Mandelbulber version '%s'", v.c_str());
    const size_t dot2 = v.find('.', dot1 + 1);
    const std::string a = v.substr(0, dot1), b = v.substr(dot1 + 1, dot2 == std::string::npos ? std::string::npos : dot2 - dot1 - 1);
    auto num = [&](const std::string &s) {
        if (s.empty() || s.find_first_not_of("0123456789") != std::string::npos)
            fail("invalid Mandelbulber version '%s'", v.c_str());
        return uint32_t(std::strtoul(s.c_str(), nullptr, 10));
    };
    const uint32_t ma = num(a), mi = num(b);
    return ma < major || (ma == major && mi < minor);
}

// Surface colour: mat1_surface_color as three 16-bit hex components.
void surface_color(const Document &d, float out[3]) {
    const std::string *v = d.value("main_parameters", "mat1_surface_color");
    if (!v) {
        for (int k = 0; k < 3; ++k) out[k] = float(50000.0 / 65535.0);
        return;
    }
    std::istringstream ss(*v);
    std::vector<std::string> parts;
    std::string tok;
    while (ss >> tok) parts.push_back(tok);
    if (parts.size() != 3) fail("expected three hexadecimal color components");
    for (int k = 0; k < 3; ++k) {
        char *end = nullptr;
        const unsigned long x = std::strtoul(parts[k].c_str(), &end, 16);
        if (parts[k].empty() || *end != 0 || x > 0xffff || parts[k][0] == '-' || parts[k][0] == '+')
            fail("invalid 16-bit color component '%s'", parts[k].c_str());
        out[k] = float(x) / 65535.0f;
    }
}

// mat1_surface_color_gradient: position / RRGGBB pairs (validated only:
// the native field does not read it).
void check_gradient(const std::string &v) {
    std::istringstream ss(v);
    std::vector<std::string> parts;
    std::string tok;
    while (ss >> tok) parts.push_back(tok);
    if (parts.size() < 2 || parts.size() % 2 != 0) fail("gradient must contain position/color pairs");
    if (parts.size() / 2 > 4096) fail("gradient contains too many color stops");
    for (size_t k = 0; k < parts.size(); k += 2) {
        if (!std::isfinite(float(parse_number(parts[k]) / 1.0) / 10000.0f)) fail("gradient position must be finite");
        if (parts[k + 1].size() != 6 || parts[k + 1].find_first_not_of("0123456789abcdefABCDEF") != std::string::npos)
            fail("invalid 8-bit RGB color '%s'", parts[k + 1].c_str());
    }
}

struct Scene {
    std::string version;
    uint32_t formula_id = 0, width = 1920, height = 1080, max_iterations = kDefaultMaxIter, projection = 0;
    uint32_t max_steps = 10000;
    bool hybrid = false, boolean_ops = false;
    Vec3 camera{0, 0, -5}, target{0, 0, 0}, camera_top{0, 0, 1};
    double fov = kDefaultFov, detail_level = 1.0, de_threshold = 0.01, detail_min = 1e-12, detail_max = 1.0;
    double smoothness = 1.0, de_factor = 1.0, abs_min = 1e-15, abs_max = 3.0, rel_min = 1e-2, rel_max = 1e4;
    double view_max = 50.0, bailout = kDefaultBailout;
    bool constant_threshold = false, iteration_threshold = false, slow_shading = false, advanced_quality = false;
    Vec3 fractal_position{0, 0, 0}, fractal_rotation{0, 0, 0}, fractal_repeat{0, 0, 0};
    bool box_fold = false, sph_fold = false;
    double box_limit = 1.0, box_value = 2.0, sph_outer = 1.0, sph_inner = 0.5;
    double ifs_scale = 2.0;
    Vec3 ifs_rotation{0, 0, 0}, ifs_offset{1, 0, 0};
    bool ifs_abs[3] = {false, false, false};
    bool ifs_enabled[kIfsVectors] = {};
    Vec3 ifs_dir[kIfsVectors];
    float surface[3] = {0, 0, 0};
};

Scene parse_scene(const std::string &src) {  // mandelbulber.rs:508-1041 (formula 10 subset)
    const Document d = parse_document(src);
    Scene s;
    s.version = d.version;
    if (d.version.compare(0, 2, "2.") != 0)
        fail("only Mandelbulber 2.x settings are currently supported, got %s", d.version.c_str());
    const char *M = "main_parameters";
    s.hybrid = boolean(d, M, "hybrid_fractal_enable", false);
    s.boolean_ops = boolean(d, M, "boolean_operators", false);
    const uint32_t repeat_from = integer(d, M, "repeat_from", 1);
    if (repeat_from < 1 || repeat_from > 9) fail("repeat_from must be 1..9");
    bool secondary = false;
    for (int i = 1; i <= kFormulaSlots; ++i) {
        const std::string n = std::to_string(i);
        const uint32_t id = integer(d, M, "formula_" + n, i == 1 ? 2u : 0u);
        const bool enabled = boolean(d, M, "fractal_enable_" + n, true);
        const uint32_t iters = integer(d, M, "formula_iterations_" + n, 1);
        if (iters < 1 || iters > 65536) fail("formula_iterations_%d must be 1..65536", i);
        const double weight = number(d, M, "formula_weight_" + n, 1.0);
        if (!(weight >= 0.0 && weight <= 1.0)) fail("formula_weight_%d must be 0..1", i);
        const uint32_t start = integer(d, M, "formula_start_iteration_" + n, 0);
        const uint32_t stop = integer(d, M, "formula_stop_iteration_" + n, 250);
        if (start > 65536 || stop > 65536) fail("formula iteration ranges must be 0..65536");
        (void)boolean(d, M, "dont_add_c_constant_" + n, false);
        (void)boolean(d, M, "check_for_bailout_" + n, true);
        const bool active = id != 0 && enabled;
        if (i == 1) {
            if (!active) fail("formula slot 1 must select an enabled formula");
            s.formula_id = id;
        } else {
            secondary = secondary || active;
        }
        const Vec3 unused0 = vec3(d, M, "formula_position_" + n, {0, 0, 0});
        const Vec3 unused1 = vec3(d, M, "formula_rotation_" + n, {0, 0, 0});
        const Vec3 unused2 = vec3(d, M, "formula_repeat_" + n, {0, 0, 0});
        (void)unused0, (void)unused1, (void)unused2;
        const double fs = number(d, M, "formula_scale_" + n, 1.0);
        if (!std::isfinite(fs) || fs == 0.0) fail("formula_scale_%d must be finite and non-zero", i);
        if (i > 1) (void)integer(d, M, "boolean_operator_" + std::to_string(i - 1), 1);
    }
    if (!(s.hybrid || s.boolean_ops || !secondary))
        fail("secondary formula slots require hybrid_fractal_enable or boolean_operators");
    if (d.value(M, "delta_DE_method")) {
        if (integer(d, M, "delta_DE_method", 0) > 2) fail("delta_DE_method must be 0..2");
    } else {
        (void)boolean(d, M, "analityc_DE_mode", true);
    }
    if (integer(d, M, "delta_DE_function", 0) > 6) fail("delta_DE_function must be 0..6");
    (void)boolean(d, M, "use_default_bailout", true);
    const double global_bailout = number(d, M, "bailout", kDefaultBailout);
    if (!std::isfinite(global_bailout) || global_bailout < 1.0) fail("global bailout must be finite and at least one");
    s.max_iterations = integer(d, M, "N", kDefaultMaxIter);
    if (s.max_iterations < 1 || s.max_iterations > 4096) fail("N must be 1..4096");
    s.width = integer(d, M, "image_width", 1920);
    s.height = integer(d, M, "image_height", 1080);
    if (s.width < 1 || s.width > 16384 || s.height < 1 || s.height > 16384) fail("image dimensions must be 1..16384");
    s.camera = vec3(d, M, "camera", {0, 0, -5});
    s.target = vec3(d, M, "target", {0, 0, 0});
    if (s.camera.x == s.target.x && s.camera.y == s.target.y && s.camera.z == s.target.z)
        fail("camera and target must differ");
    s.camera_top = vec3(d, M, "camera_top", {0, 0, 1});
    const std::string *persp = d.value(M, "perspective_type");
    const std::string p = persp ? *persp : "three_point";
    if (p == "three_point" || p == "0") s.projection = 0;
    else if (p == "fish_eye" || p == "1") s.projection = 1;
    else if (p == "equirectangular" || p == "2") s.projection = 2;
    else if (p == "fish_eye_cut" || p == "3") s.projection = 3;
    else fail("unsupported perspective_type %s", p.c_str());
    const double stored_fov = number(d, M, "fov", kDefaultFov);
    s.fov = stored_fov;
    if (version_before(d.version, 2, 21)) {  // mandelbulber.rs:361-379
        const double legacy = stored_fov == kDefaultFov ? 1.0 : stored_fov;
        if (s.projection == 0) s.fov = to_degrees(std::atan(legacy * 0.5)) * 2.0;
        else if (s.projection == 1 || s.projection == 3) s.fov = legacy * 180.0;
        else s.fov = legacy * 360.0;
    }
    const double max_fov = s.projection == 0 ? 180.0 : 720.0;
    if (!(s.fov > 0.0 && s.fov <= max_fov))
        fail("field of view must be positive and at most %g degrees for this projection", max_fov);
    s.detail_level = number(d, M, "detail_level", 1.0);
    if (!(s.detail_level > 0.0)) fail("detail_level must be positive");
    s.constant_threshold = boolean(d, M, "constant_DE_threshold", false);
    s.de_threshold = number(d, M, "DE_thresh", 0.01);
    s.iteration_threshold = boolean(d, M, "iteration_threshold_mode", false);
    s.detail_min = number(d, M, "detail_size_min", 1.0e-12);
    s.detail_max = number(d, M, "detail_size_max", 1.0);
    s.smoothness = number(d, M, "smoothness", 1.0);
    s.slow_shading = boolean(d, M, "slow_shading", false);
    s.fractal_position = vec3(d, M, "fractal_position", {0, 0, 0});
    s.fractal_rotation = vec3(d, M, "fractal_rotation", {0, 0, 0});
    s.fractal_repeat = vec3(d, M, "repeat", {0, 0, 0});
    s.de_factor = number(d, M, "DE_factor", 1.0);
    s.advanced_quality = boolean(d, M, "advanced_quality", false);
    s.abs_min = number(d, M, "abs_min_marching_step", 1.0e-15);
    s.abs_max = number(d, M, "abs_max_marching_step", 3.0);
    s.rel_min = number(d, M, "rel_min_marching_step", 1.0e-2);
    s.rel_max = number(d, M, "rel_max_marching_step", 1.0e4);
    s.max_steps = integer(d, M, "max_raymarching_steps", 10000);
    if (!(s.de_threshold > 0.0 && s.detail_min > 0.0 && s.detail_max >= s.detail_min && s.smoothness > 0.0 &&
          s.de_factor > 0.0 && s.abs_min > 0.0 && s.abs_max >= s.abs_min && s.rel_min > 0.0 && s.rel_max >= s.rel_min))
        fail("distance-threshold parameters must be positive and ordered");
    if (s.max_steps < 1 || s.max_steps > 100000000) fail("max_raymarching_steps must be 1..100000000");
    s.view_max = number(d, M, "view_distance_max", 50.0);
    if (!(s.view_max > 0.0)) fail("view_distance_max must be positive");
    const char *F = "fractal_1";
    s.ifs_scale = number(d, F, "IFS_scale", 2.0);
    s.ifs_rotation = vec3(d, F, "IFS_rotation", {0, 0, 0});
    (void)boolean(d, F, "IFS_rotation_enabled", false);
    s.ifs_offset = vec3(d, F, "IFS_offset", {1, 0, 0});
    (void)boolean(d, F, "IFS_edge_enabled", false);
    (void)boolean(d, F, "IFS_menger_sponge_mode", false);
    s.ifs_abs[0] = boolean(d, F, "IFS_abs_x", false);
    s.ifs_abs[1] = boolean(d, F, "IFS_abs_y", false);
    s.ifs_abs[2] = boolean(d, F, "IFS_abs_z", false);
    for (int k = 0; k < kIfsVectors; ++k) {
        const std::string n = std::to_string(k);
        s.ifs_enabled[k] = boolean(d, F, "IFS_enabled_" + n, false);
        s.ifs_dir[k] = normalized(vec3(d, F, "IFS_direction_" + n, {1, 0, 0}));
        (void)vec3(d, F, "IFS_rotations_" + n, {0, 0, 0});
        (void)number(d, F, "IFS_distance_" + n, 0.0);
        (void)number(d, F, "IFS_intensity_" + n, 1.0);
    }
    if (s.formula_id == 10 && s.ifs_scale == 0.0) fail("IFS_scale must be non-zero");
    (void)number(d, F, "mandelbox_folding_min_radius", 0.5);
    static const char *fold_names[2] = {"neg", "pos"};
    for (const char *fn : fold_names)
        for (int a = 1; a <= 3; ++a) (void)vec3(d, F, std::string("mandelbox_rotation_") + fn + "_" + std::to_string(a), {0, 0, 0});
    surface_color(d, s.surface);
    const std::string *grad = d.value(M, "mat1_surface_color_gradient");
    check_gradient(grad ? *grad : std::string("0 fd6029 1000 698403 2000 fff59b 3000 f5bd22 4000 0b5e87"));
    (void)boolean(d, M, "mat1_use_colors_from_palette", true);
    (void)boolean(d, M, "mat1_surface_gradient_enable", true);
    (void)number(d, M, "mat1_coloring_speed", 1.0);
    (void)number(d, M, "mat1_coloring_palette_offset", 0.0);
    s.bailout = s.formula_id == 3 ? 10.0 : kDefaultBailout;
    s.box_fold = boolean(d, M, "box_folding", false);
    s.box_limit = number(d, M, "box_folding_limit", 1.0);
    s.box_value = number(d, M, "box_folding_value", 2.0);
    s.sph_fold = boolean(d, M, "spherical_folding", false);
    s.sph_outer = number(d, M, "spherical_folding_outer", 1.0);
    s.sph_inner = number(d, M, "spherical_folding_inner", 0.5);
    return s;
}

// Mandelbulber's (x, y, z) to the renderer's (x, z, y) (mandelbulber.rs:2483-2489).
Vec3 map_point(Vec3 p) { return {p.x, p.z, p.y}; }

double roll_of(Vec3 dir, Vec3 up, float yaw_f, float pitch_f) {  // mandelbulber.rs:2491-2503
    const Vec3 f = normalized(dir);
    Vec3 pu = sub(up, mul(f, dot(up, f)));
    const double m = vlen(pu);
    pu = (m > 0.0 && std::isfinite(m)) ? mul(pu, 1.0 / m) : Vec3{0, 1, 0};
    const double yaw = double(yaw_f), pitch = double(pitch_f);
    const Vec3 right{std::cos(yaw), 0.0, -std::sin(yaw)};
    const Vec3 unrolled{-std::sin(yaw) * std::sin(pitch), std::cos(pitch), -std::cos(yaw) * std::sin(pitch)};
    return std::atan2(-dot(pu, right), dot(pu, unrolled));
}

void apply_to_config(const Scene &s, FptRenderConfig &c) {  // mandelbulber.rs:1177-1344 (formula 10)
    c.sdf_id = FPT_SDF_MANDELBULBER;
    c.width = s.width;
    c.height = s.height;
    const Vec3 cam = mul(map_point(s.camera), kWorldScale), tgt = mul(map_point(s.target), kWorldScale);
    const Vec3 dir = sub(tgt, cam);
    c.camera_position[0] = float(cam.x);
    c.camera_position[1] = float(cam.y);
    c.camera_position[2] = float(cam.z);
    const double horizontal = std::hypot(dir.x, dir.z);
    c.camera_yaw_pitch[0] = float(std::atan2(dir.x, dir.z));
    c.camera_yaw_pitch[1] = float(std::atan2(dir.y, horizontal));
    c.camera_roll = float(roll_of(dir, map_point(s.camera_top), c.camera_yaw_pitch[0], c.camera_yaw_pitch[1]));
    // Mandelbulber spans [-0.5, 0.5] on the image plane times 2 tan(fov / 2);
    // this renderer's focal length is 1 / tan(fov / 2).
    c.camera_fov = s.projection == 0 ? float(to_degrees(2.0 * std::atan(2.0 * std::tan(to_radians(s.fov) * 0.5))))
                                     : float(s.fov);
    c.camera_dof = 0.0f;
    c.focus_distance = float(vlen(dir) * kWorldScale);
    const double internal_fov = s.projection == 0   ? 2.0 * std::tan(to_radians(s.fov) * 0.5)
                                : s.projection == 2 ? to_radians(s.fov) * 0.5
                                                    : to_radians(s.fov);
    const double threshold_scale = internal_fov / double(s.height) / (s.iteration_threshold ? 1.0 : s.detail_level);
    double distance_threshold;
    if (s.constant_threshold && !s.iteration_threshold) {
        distance_threshold = s.de_threshold * kWorldScale;
    } else {
        const double lo = s.detail_min * kWorldScale, hi = s.detail_max * kWorldScale;
        distance_threshold = std::min(std::max(vlen(dir) * threshold_scale, lo), hi);
    }
    c.render[1] = float(s.max_steps);
    c.render[2] = float(distance_threshold);
    c.render[3] = float(distance_threshold);
    c.render[4] = float(s.view_max * kWorldScale);
    c.render[5] = 0.0f;
    c.render[7] = s.slow_shading ? float(internal_fov / double(s.height)) : 0.0f;
    const float world[7] = {3.0f, 1.0f, 125.0f, 30.0f, 0.45f, 1.0f, 1.0f};
    std::memcpy(c.world, world, sizeof(world));
    const float one[3] = {0.72f, 0.80f, 0.95f};
    std::memcpy(c.world_one_color, one, sizeof(one));
    const float bg[6] = {0.08f, 0.11f, 0.16f, 0.008f, 0.012f, 0.02f};
    std::memcpy(c.background_gradient, bg, sizeof(bg));
    const float sun[5] = {1.0f, 130.0f, 32.0f, 1.6f, 0.055f};
    std::memcpy(c.sun, sun, sizeof(sun));
    const float sc[3] = {1.0f, 0.82f, 0.68f};
    std::memcpy(c.sun_color, sc, sizeof(sc));
    c.post[1] = 1.0f;
    c.set_values[P_WORLD_SCALE] = float(kWorldScale);
    c.set_values[P_MAX_ITER] = float(s.max_iterations);
    c.set_values[P_BAILOUT] = float(s.bailout);
    c.vset_values[V_BOX_FOLD] = s.box_fold ? 1.0f : 0.0f;
    c.vset_values[V_BOX_LIMIT] = float(s.box_limit);
    c.vset_values[V_BOX_VALUE] = float(s.box_value);
    c.vset_values[V_SPH_FOLD] = s.sph_fold ? 1.0f : 0.0f;
    c.vset_values[V_SPH_OUTER] = float(s.sph_outer);
    c.vset_values[V_SPH_INNER] = float(s.sph_inner);
    c.vset_values[V_ADV_QUALITY] = s.advanced_quality ? 1.0f : 0.0f;
    c.vset_values[V_ABS_MIN] = float(s.abs_min * kWorldScale);
    c.vset_values[V_ABS_MAX] = float(s.abs_max * kWorldScale);
    c.vset_values[V_REL_MIN] = float(s.rel_min);
    c.vset_values[V_REL_MAX] = float(s.rel_max);
    c.vset_values[V_DE_FACTOR] = float(s.de_factor);
    c.vset_values[V_SMOOTH] = float(s.smoothness);
    c.vset_values[V_DYN_THRESH] = s.iteration_threshold ? kIterationThresholdMode : (s.constant_threshold ? 0.0f : 1.0f);
    c.vset_values[V_THRESH_SCALE] = float(threshold_scale);
    c.vset_values[V_CONST_THRESH] = float(s.de_threshold * kWorldScale);
    c.vset_values[V_MIN_THRESH] = float(s.detail_min * kWorldScale);
    c.vset_values[V_MAX_THRESH] = float(s.detail_max * kWorldScale);
    c.vset_values[V_PROJECTION] = float(s.projection);
    const double pos[3] = {s.fractal_position.x, s.fractal_position.y, s.fractal_position.z};
    const double rot[3] = {s.fractal_rotation.x, s.fractal_rotation.y, s.fractal_rotation.z};
    const double rep[3] = {s.fractal_repeat.x, s.fractal_repeat.y, s.fractal_repeat.z};
    for (int k = 0; k < 3; ++k) {
        c.vset_values[V_POSITION + k] = float(pos[k]);
        c.vset_values[V_ROTATION + k] = float(to_radians(rot[k]));
        c.vset_values[V_REPEAT + k] = float(rep[k]);
    }
    c.vset_values[V_DETAIL] = float(s.detail_level);
    c.vset_values[V_REFINEMENT] = 0.0f;  // interactive preview controls (environment-driven there): off
    c.vset_values[V_SPATIAL] = 0.0f;
    c.set_values[P_IFS_SCALE] = float(s.ifs_scale);
    uint32_t abs_mask = 0, en_mask = 0;
    for (int k = 0; k < 3; ++k) abs_mask |= uint32_t(s.ifs_abs[k]) << k;
    for (int k = 0; k < kIfsVectors; ++k) en_mask |= uint32_t(s.ifs_enabled[k]) << k;
    c.set_values[P_ABS_MASK] = float(abs_mask);
    c.set_values[P_ENABLED_MASK] = float(en_mask);
    c.set_values[P_OFFSET + 0] = float(s.ifs_offset.x);
    c.set_values[P_OFFSET + 1] = float(s.ifs_offset.y);
    c.set_values[P_OFFSET + 2] = float(s.ifs_offset.z);
    c.set_values[P_ROTATION + 0] = float(to_radians(s.ifs_rotation.x));
    c.set_values[P_ROTATION + 1] = float(to_radians(s.ifs_rotation.y));
    c.set_values[P_ROTATION + 2] = float(to_radians(s.ifs_rotation.z));
    for (int k = 0; k < kIfsVectors; ++k) {
        c.set_values[P_DIRECTION + 3 * k + 0] = float(s.ifs_dir[k].x);
        c.set_values[P_DIRECTION + 3 * k + 1] = float(s.ifs_dir[k].y);
        c.set_values[P_DIRECTION + 3 * k + 2] = float(s.ifs_dir[k].z);
    }
    c.set_values[P_FORMULA_ID] = float(s.formula_id);
    c.fractal_style_mode = 2;  // formula 10 is not an analytic-DE formula
    c.fractal_style[1] = 0.0f;
    c.fractal_style[2] = 0.0f;
    c.fractal_style[3] = 1.0f;
    c.fractal_style[4] = 0.72f;
    c.fractal_style[5] = 0.08f;
    c.fractal_style[6] = 0.0f;
    std::memcpy(&c.fractal_style[8], s.surface, sizeof(s.surface));
}

}  // namespace

// load_mandelbulber_scene_config (scene.rs:1978-2037) for formula 10.
bool fract_load(const std::string &path, const FptSceneOptions &opt, FptRenderConfig &c, std::string &output_name,
                char *error, size_t error_len) {
    auto set = [&](const std::string &m) {
        if (error && error_len) std::snprintf(error, error_len, "%s", m.c_str());
        return false;
    };
    if (opt.renderer_backend != FPT_RENDERER_SDF) return set("Mandelbulber scenes currently support only the direct sdf renderer");
    if (opt.topology_specialization || opt.flat_union || opt.typed_soa)
        return set("typed-program compiler options do not apply to Mandelbulber iterative-DE scenes");
    std::ifstream in(path, std::ios::binary);
    if (!in) return set("read Mandelbulber scene " + path + ": cannot open");
    std::ostringstream ss;
    ss << in.rdbuf();
    try {
        Scene s = parse_scene(ss.str());
        if (opt.width) s.width = opt.width;
        if (opt.height) s.height = opt.height;
        if (s.hybrid || s.formula_id != 10)
            return set("formula ID " + std::to_string(s.formula_id) +
                       " needs the Mandelbulber2 formula source; only the native Kaleidoscopic IFS (formula 10, no hybrid) is supported");
        scene_default_config(c);
        apply_to_config(s, c);
        c.preview = opt.preview;
        c.sdf_profile = opt.sdf_profile;
        c.width = s.width;
        c.height = s.height;
        c.samples = opt.samples ? opt.samples : 64u;
        if (c.samples < 1u || c.samples > 512u) return set("scene samples must be 1..512");
    } catch (const Fail &f) {
        return set("import Mandelbulber scene " + path + ": " + f.msg);
    }
    const size_t slash = path.find_last_of('/');
    std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = name.find_last_of('.');
    if (dot != std::string::npos && dot > 0) name = name.substr(0, dot);
    output_name = (name.empty() ? std::string("mandelbulber") : name) + ".png";
    return true;
}

}  // namespace fpt
