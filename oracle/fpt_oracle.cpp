// ORACLE — test infrastructure only (see fpt_oracle.h).
//
// CPU restatement of the reference's SDF hot path.  Every function cites the
// reference lines it follows.  Abbreviations: SM = shaders/Shaders.metal,
// BR = src/metal_bridge.mm, ORB = src/mandelbulber/orbit.rs.
//
// Pins (checks that a plausible mistake would fail), in tests/:
//   hash13                 integer KAT vectors computed by hand from the
//                          MurmurHash2 mixing constants; float mapping range.
//   mandelbulb step        ORB:1876-1888 (p=2 from (1,2,3): dr = 2r+1,
//                          |z| = r^2) and the triplex closed form at p=8.
//   mandelbox step         ORB:1915-1960 (scale -1.5, minR2 .25, fixedR2 1
//                          from (.1,.2,.3) -> (-.6,-1.2,-1.8), dr 7).
//   leaf orbit / DE        brute-force iteration in numpy (tests), and the
//                          exact closed forms for bailout on iteration 1.
//   program interpreter    closed-form sphere/box/plane/CSG distances of
//                          scenes/benchmarks/Exact_*.json.
//   program gradient       FP64 central-difference limit (brute force).
//   lowered evaluators     (tests/test_lowered_programs.py) the interpreter
//                          of the same program on every bundled scene each
//                          lowering accepts (FP64 distances bit-identical);
//                          a plane/sphere tie for the typed table's
//                          earlier-instruction rule; union, intersection
//                          and subtraction of a canonical list in numpy.
//   march                  invariants: hit => 0.99|d| < eps at the last step;
//                          analytic hit t for a sphere at the origin.
//   render                 FP32 twin vs FP64 (agreement), sphere scene
//                          direct-light closed form at the image centre.
//   Mandelbulber scenes    (tests/test_mandelbulber_render_oracle.py) a KIFS
//                          whose field is an exact sphere distance: march
//                          end points in the refinement band, radial
//                          normals (central, tetrahedral, slow shading), the
//                          iteration-limit interior rule; the projections'
//                          defining angles; the (0, 1) pixel offset.
//                          normal_mandelbulber's log-orbit-radius fallback
//                          (six equal central samples): parity unpinned.
//
// The RNG (SM:541-576) keys on the IEEE bit patterns of float32 values, so
// every RNG key and the pixel uv are computed in float32 with no contraction
// in both precision modes; geometry runs in T.
#include "fpt_oracle.h"

#include <algorithm>
#include <atomic>
#include <utility>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265359;  // SM:4 (the reference's rounded pi)
constexpr float kPiF = 3.14159265359f;

// ---------------------------------------------------------------------------
// Work counters (SM:7650-7704 categories).
// ---------------------------------------------------------------------------
struct Counters {
    uint64_t distance_evals = 0, orbit_iterations = 0, primary_steps = 0,
             secondary_steps = 0, shadow_steps = 0, normal_evals = 0, paths = 0;
    // march steps whose distance exceeded 0.02 (far field) and their orbit
    // iterations: the work a far-field step proposer could replace
    uint64_t far_steps = 0, far_orbit_iterations = 0;
    void add(const Counters &o) {
        far_steps += o.far_steps;
        far_orbit_iterations += o.far_orbit_iterations;
        distance_evals += o.distance_evals;
        orbit_iterations += o.orbit_iterations;
        primary_steps += o.primary_steps;
        secondary_steps += o.secondary_steps;
        shadow_steps += o.shadow_steps;
        normal_evals += o.normal_evals;
        paths += o.paths;
    }
};
thread_local Counters *tl_counters = nullptr;
Counters g_last;

inline void count_orbit(uint64_t n) {
    if (tl_counters) tl_counters->orbit_iterations += n;
}

// ---------------------------------------------------------------------------
// Scalar helpers generic over T (double, float) and Dual<T>.
// ---------------------------------------------------------------------------
template <class T>
struct Dual {  // forward-mode value + gradient wrt the 3 point coordinates
    T v, d0, d1, d2;
    Dual() : v(0), d0(0), d1(0), d2(0) {}
    Dual(T x) : v(x), d0(0), d1(0), d2(0) {}  // NOLINT (implicit constant)
    Dual(T x, T a, T b, T c) : v(x), d0(a), d1(b), d2(c) {}
};
template <class T> Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2}; }
template <class T> Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2}; }
template <class T> Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d0, -a.d1, -a.d2}; }
template <class T> Dual<T> operator*(Dual<T> a, Dual<T> b) {
    return {a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2};
}
template <class T> Dual<T> operator/(Dual<T> a, Dual<T> b) {
    T q = a.v / b.v;
    return {q, (a.d0 - q * b.d0) / b.v, (a.d1 - q * b.d1) / b.v, (a.d2 - q * b.d2) / b.v};
}
template <class T> Dual<T> scale_d(Dual<T> a, T k) { return {a.v * k, a.d0 * k, a.d1 * k, a.d2 * k}; }

template <class T> inline T val(T x) { return x; }
template <class T> inline T val(Dual<T> x) { return x.v; }

template <class T> inline T s_sqrt(T x) { return std::sqrt(x); }
template <class T> Dual<T> s_sqrt(Dual<T> x) {
    T r = std::sqrt(x.v);
    T k = T(0.5) / r;
    return {r, x.d0 * k, x.d1 * k, x.d2 * k};
}
template <class T> inline T s_sin(T x) { return std::sin(x); }
template <class T> Dual<T> s_sin(Dual<T> x) { T c = std::cos(x.v); return {std::sin(x.v), x.d0 * c, x.d1 * c, x.d2 * c}; }
template <class T> inline T s_cos(T x) { return std::cos(x); }
template <class T> Dual<T> s_cos(Dual<T> x) { T s = -std::sin(x.v); return {std::cos(x.v), x.d0 * s, x.d1 * s, x.d2 * s}; }
template <class T> inline T s_asin(T x) { return std::asin(x); }
template <class T> Dual<T> s_asin(Dual<T> x) {
    T k = T(1) / std::sqrt(T(1) - x.v * x.v);
    return {std::asin(x.v), x.d0 * k, x.d1 * k, x.d2 * k};
}
template <class T> inline T s_atan2(T y, T x) { return std::atan2(y, x); }
template <class T> Dual<T> s_atan2(Dual<T> y, Dual<T> x) {
    T den = x.v * x.v + y.v * y.v;
    T a = x.v / den, b = -y.v / den;  // d/dy, d/dx
    return {std::atan2(y.v, x.v), a * y.d0 + b * x.d0, a * y.d1 + b * x.d1, a * y.d2 + b * x.d2};
}
template <class T> inline T s_log(T x) { return std::log(x); }
template <class T> Dual<T> s_log(Dual<T> x) { T k = T(1) / x.v; return {std::log(x.v), x.d0 * k, x.d1 * k, x.d2 * k}; }
template <class T> inline T s_pow(T x, T e) { return std::pow(x, e); }
template <class T> Dual<T> s_pow(Dual<T> x, T e) {
    T p = std::pow(x.v, e);
    T k = e * std::pow(x.v, e - T(1));
    return {p, x.d0 * k, x.d1 * k, x.d2 * k};
}
template <class T> inline T s_abs(T x) { return std::fabs(x); }
template <class T> Dual<T> s_abs(Dual<T> x) { return x.v < T(0) ? -x : x; }
template <class T> inline T s_min(T a, T b) { return b < a ? b : a; }       // fmin semantics for finite
template <class T> inline T s_max(T a, T b) { return a < b ? b : a; }
template <class T> Dual<T> s_min(Dual<T> a, Dual<T> b) { return b.v < a.v ? b : a; }
template <class T> Dual<T> s_max(Dual<T> a, Dual<T> b) { return a.v < b.v ? b : a; }

// Metal min/max/clamp on floats follow fmin/fmax (NaN-ignoring).  Plain
// scalars use std::fmin/fmax to match.
inline double m_min(double a, double b) { return std::fmin(a, b); }
inline double m_max(double a, double b) { return std::fmax(a, b); }
inline float m_min(float a, float b) { return std::fmin(a, b); }
inline float m_max(float a, float b) { return std::fmax(a, b); }
template <class T> inline T m_clamp(T x, T lo, T hi) { return m_min(m_max(x, lo), hi); }
template <class T> inline T m_sign(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }
template <class T> inline T m_fract(T x) { return x - std::floor(x); }
template <class T> inline T m_mix(T a, T b, T t) { return a + (b - a) * t; }

template <class S>
struct V3 {
    S x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(S a, S b, S c) : x(a), y(b), z(c) {}
};
template <class S> V3<S> operator+(V3<S> a, V3<S> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class S> V3<S> operator-(V3<S> a, V3<S> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class S> V3<S> operator-(V3<S> a) { return {-a.x, -a.y, -a.z}; }
template <class S> V3<S> operator*(V3<S> a, V3<S> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class S, class K> V3<S> operator*(V3<S> a, K k) { return {a.x * S(k), a.y * S(k), a.z * S(k)}; }
template <class S, class K> V3<S> operator/(V3<S> a, K k) { return {a.x / S(k), a.y / S(k), a.z / S(k)}; }
template <class S> S dot(V3<S> a, V3<S> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> S length(V3<S> a) { return s_sqrt(dot(a, a)); }
template <class S> V3<S> normalize(V3<S> a) { return a / length(a); }
template <class S> V3<S> cross(V3<S> a, V3<S> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class S> V3<S> vabs(V3<S> a) { return {s_abs(a.x), s_abs(a.y), s_abs(a.z)}; }
template <class S> V3<S> vmix(V3<S> a, V3<S> b, S t) { return a + (b - a) * t; }
template <class S> V3<S> vreflect(V3<S> i, V3<S> n) { return i - n * (S(2) * dot(n, i)); }
template <class S> V3<S> vrefract(V3<S> i, V3<S> n, S eta) {
    S ni = dot(n, i);
    S k = S(1) - eta * eta * (S(1) - ni * ni);
    if (k < S(0)) return V3<S>();
    return i * eta - n * (eta * ni + std::sqrt(k));
}
template <class S> bool vfinite(V3<S> a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

template <class T> V3<T> load3(const float *v) { return {T(v[0]), T(v[1]), T(v[2])}; }

// ---------------------------------------------------------------------------
// RNG (SM:541-576).  Integer hash of float32 bit patterns.
// ---------------------------------------------------------------------------
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

uint32_t hash13bits(float x, float y, float z) {
    uint32_t h = 1190494759u;
    h = (h ^ fbits(x)) * 0x5bd1e995u;
    h = (h ^ fbits(y)) * 0x5bd1e995u;
    h = (h ^ fbits(z)) * 0x5bd1e995u;
    h ^= h >> 13u;
    h *= 0x5bd1e995u;
    h ^= h >> 15u;
    return h;
}
float hash13(float x, float y, float z) {
    return float(hash13bits(x, y, z) & 0x00ffffffu) / float(0x01000000u);
}

struct UV { float x, y; };

// randomVector (SM:557-568): cosine-weighted hemisphere about `normal`.
template <class T>
V3<T> random_vector(V3<T> normal, UV xy, float seed) {
    volatile float k0 = seed * 2.0f;  // volatile: keep float32 rounding, no contraction
    float s0 = k0 + 0.0f, s1 = k0 + 1.0f;
    T h1 = T(hash13(xy.x, xy.y, s0));
    T h2 = T(hash13(xy.x, xy.y, s1));
    V3<T> n = normalize(normal);
    V3<T> uu = normalize(cross(n, V3<T>(T(0), T(1), T(1))));
    V3<T> vv = cross(uu, n);
    T ra = std::sqrt(h2);
    T rx = ra * std::cos(T(kPi) * T(2) * h1);
    T ry = ra * std::sin(T(kPi) * T(2) * h1);
    T rz = std::sqrt(m_max(T(1) - h2, T(0)));
    return normalize(uu * rx + vv * ry + n * rz);
}

// randomPoint (SM:570-576).
template <class T>
void random_point(T power, UV xy, float seed, T &ox, T &oy) {
    volatile float k0 = seed * 2.0f;
    float s0 = k0 + 0.0f, s1 = k0 + 1.0f;
    T h1 = T(hash13(xy.x, xy.y, s0));
    T h2 = T(hash13(xy.x, xy.y, s1));
    T r = T(2) * h1 * T(kPi);
    T r2 = T(2) * h2 * T(kPi);
    ox = std::cos(r) * std::sqrt(r2) * power;
    oy = std::sin(r) * std::sqrt(r2) * power;
}

// float32 a*b + c with the product rounded first (no contraction).
inline float mul_add_f32(float a, float b, float c) {
    volatile float p = a * b;
    return p + c;
}

// ---------------------------------------------------------------------------
// Camera (SM:460-471, 6847-6855, 5940-5947).
// ---------------------------------------------------------------------------
template <class T>
V3<T> rotate_camera(V3<T> v, T yaw, T pitch) {
    v = V3<T>(v.x, v.z * std::sin(pitch) + v.y * std::cos(pitch), v.z * std::cos(pitch) - v.y * std::sin(pitch));
    v = V3<T>(v.x * std::cos(yaw) + v.z * std::sin(yaw), v.y, -v.x * std::sin(yaw) + v.z * std::cos(yaw));
    return v;
}
template <class T>
V3<T> rotate_camera_roll(V3<T> v, T yaw, T pitch, T roll) {
    T s = std::sin(roll), c = std::cos(roll);
    V3<T> r(c * v.x - s * v.y, s * v.x + c * v.y, v.z);
    return rotate_camera(r, yaw, pitch);
}

// sdfScreenUv (SM:6847-6855) in float32, as the GPU computes it: pixel
// centres, or the pixel offset (0, 1) for Mandelbulber scenes.
UV screen_uv(uint32_t x, uint32_t y, const FptRenderConfig &cfg) {
    const bool mb = cfg.sdf_id == FPT_SDF_MANDELBULBER;
    volatile float fw = float(cfg.width), fh = float(cfg.height);
    volatile float ux = (float(x) + (mb ? 0.0f : 0.5f)) / fw;
    volatile float uy = (float(y) + (mb ? 1.0f : 0.5f)) / fh;
    float u0 = ux - 0.5f, u1 = uy - 0.5f;
    volatile float aspect = fw / fh;
    volatile float u0a = u0 * aspect;
    return {u0a, u1};
}

template <class T>
V3<T> camera_dir(T sx, T sy, const FptRenderConfig &cfg) {
    T focal = T(1) / std::tan(T(cfg.camera_fov) / T(2) * T(kPi) / T(180));
    return rotate_camera_roll(normalize(V3<T>(sx, sy, focal)), T(cfg.camera_yaw_pitch[0]),
                              T(cfg.camera_yaw_pitch[1]), T(cfg.camera_roll));
}

// The camera ray of renderPath (SM:5938-5949): Mandelbulber scenes use
// mandelbulberCameraRay (SM:482-511) with vset_values[107] as the projection
// (0 perspective, 1 fish-eye, 2 equirectangular, 3 fish-eye cut at 180
// degrees, whose outside is not seen: false).
template <class T>
bool camera_ray(T sx, T sy, const FptRenderConfig &cfg, V3<T> &dr) {
    if (cfg.sdf_id != FPT_SDF_MANDELBULBER) {
        dr = camera_dir<T>(sx, sy, cfg);
        return true;
    }
    const float proj = cfg.vset_values[107];
    const T fov = T(cfg.camera_fov) * T(kPi) / T(180);
    const T radius = std::sqrt(sx * sx + sy * sy);
    if (proj >= 2.5f && proj <= 3.5f && !(radius <= T(0.5) * T(kPi) / m_max(fov, T(1.0e-6f)))) return false;
    V3<T> d;
    if ((proj > 0.5f && proj < 1.5f) || (proj > 2.5f && proj < 3.5f)) {
        if (radius <= T(1.0e-8f)) {
            d = V3<T>(T(0), T(0), T(1));
        } else {
            const T k = std::sin(radius * fov) / radius;
            d = V3<T>(sx * k, sy * k, std::cos(radius * fov));
        }
    } else if (proj > 1.5f && proj < 2.5f) {
        const T aspect = T(float(cfg.width) / float(cfg.height));
        const T ax = sx * (T(2) / aspect) * (T(0.5) * fov), ay = sy * (T(0.5) * fov);
        d = V3<T>(std::sin(ax) * std::cos(ay), std::sin(ay), std::cos(ax) * std::cos(ay));
    } else {
        d = normalize(V3<T>(sx, sy, T(1) / std::tan(T(0.5) * fov)));
    }
    dr = rotate_camera_roll(normalize(d), T(cfg.camera_yaw_pitch[0]), T(cfg.camera_yaw_pitch[1]),
                            T(cfg.camera_roll));
    return true;
}

// ---------------------------------------------------------------------------
// Fractal leaves (new opcodes; semantics from ORB and SM).
// ---------------------------------------------------------------------------
template <class S>
struct LeafResult {
    S d;
    double radius = 0, derivative = 0, trap = 0;
    int iterations = 0;
    bool escaped = false;
    V3<double> z;
};


// Orbit radius |z|.  In float32, |z|^2 overflows once |z| > 1.8e19, which the
// Mandelbulb reaches after one iteration from |p| ~ 235; the scaled form
// keeps r finite there (in double it equals sqrt(dot(z, z)) to 1 ulp).
template <class S>
S orbit_radius(V3<S> z) {
    S r2 = dot(z, z);
    if (std::isfinite(val(r2))) return s_sqrt(r2);
    auto T_abs = [](S v) { return s_abs(v); };
    S m = s_max(T_abs(z.x), s_max(T_abs(z.y), T_abs(z.z)));
    if (!std::isfinite(val(m))) return r2;
    V3<S> q(z.x / m, z.y / m, z.z / m);
    return m * s_sqrt(dot(q, q));
}

// Mandelbulb kernel step (ORB:766-778) on (z, r, dr).
template <class S, class T>
void mandelbulb_kernel(V3<S> &z, S r, S &dr, T power) {
    S theta0 = s_asin(z.z / r);
    S phi0 = s_atan2(z.y, z.x);
    S rp = s_pow(r, power - T(1));
    S theta = theta0 * S(power);
    S phi = phi0 * S(power);
    dr = rp * dr * S(power) + S(1);
    rp = rp * r;
    S ct = s_cos(theta);
    z = V3<S>(ct * s_cos(phi) * rp, ct * s_sin(phi) * rp, s_sin(theta) * rp);
}

// Mandelbulb leaf: orbit loop (ORB:530-585) + Logarithmic finalizer
// (ORB:654-660), add_c on (mandelbulber.rs:964, 1421-1440).
template <class S, class T>
LeafResult<S> leaf_mandelbulb(V3<S> p, const float data[4]) {
    T power = T(data[0]);
    int iters = int(data[1]);
    T bailout = T(data[2]);
    V3<S> z = p, c = p;
    S r = length(z);
    S dr = S(1);
    double trap = double(val(r));
    int done = 0;
    bool escaped = false;
    for (int i = 0; i < iters; ++i) {
        V3<S> old = z;
        mandelbulb_kernel<S, T>(z, r, dr, power);
        z = z + c;
        r = orbit_radius(z);
        done = i + 1;
        if (!std::isfinite(val(r))) {
            z = old;
            r = orbit_radius(z);
            break;
        }
        trap = std::fmin(trap, double(val(r)));
        if (val(r) > bailout) { escaped = true; break; }
    }
    count_orbit(uint64_t(done));
    LeafResult<S> out;
    if (!(val(dr) > T(0))) {
        out.d = r;
    } else if (val(r) > T(1)) {
        out.d = S(T(0.5)) * r * s_log(r) / dr;
    } else {
        out.d = S(T(0));
    }
    out.radius = double(val(r));
    out.derivative = double(val(dr));
    out.iterations = done;
    out.escaped = escaped;
    out.z = V3<double>(double(val(z.x)), double(val(z.y)), double(val(z.z)));
    out.trap = trap;
    return out;
}

// Mandelbox kernel step (ORB:1636-1687; fold limit 1, value 2, no offset,
// no rotation) on (z, dr).
template <class S, class T>
void mandelbox_kernel(V3<S> &z, S &dr, T scale, T min_r2, T fixed_r2) {
    S *c[3] = {&z.x, &z.y, &z.z};
    for (int a = 0; a < 3; ++a) {
        T v = val(*c[a]);
        if (std::fabs(v) > T(1)) {
            T sgn = v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0));
            *c[a] = S(sgn * T(2)) - *c[a];
        }
    }
    S r2 = dot(z, z);
    if (val(r2) < min_r2) {
        T f = fixed_r2 / min_r2;
        z = z * f;
        dr = dr * S(f);
    } else if (val(r2) < fixed_r2) {
        S f = S(fixed_r2) / r2;
        z = z * f;
        dr = dr * f;
    }
    z = z * scale;
    dr = dr * S(std::fabs(scale)) + S(1);
}

// Mandelbox leaf: orbit loop (ORB:530-585), bailout 100
// (mandelbulber.rs:15), Linear finalizer offset 0 (ORB:661).
template <class S, class T>
LeafResult<S> leaf_mandelbox(V3<S> p, const float data[4]) {
    T scale = T(data[0]);
    int iters = int(data[1]);
    T min_r = T(data[2]), fixed_r = T(data[3]);
    T min_r2 = min_r * min_r, fixed_r2 = fixed_r * fixed_r;
    V3<S> z = p, c = p;
    S r = length(z);
    S dr = S(1);
    double trap = double(val(r));
    int done = 0;
    bool escaped = false;
    for (int i = 0; i < iters; ++i) {
        V3<S> old = z;
        mandelbox_kernel<S, T>(z, dr, scale, min_r2, fixed_r2);
        z = z + c;
        r = orbit_radius(z);
        done = i + 1;
        if (!std::isfinite(val(r))) {
            z = old;
            r = orbit_radius(z);
            break;
        }
        trap = std::fmin(trap, double(val(r)));
        if (val(r) > T(100)) { escaped = true; break; }
    }
    count_orbit(uint64_t(done));
    LeafResult<S> out;
    out.d = !(val(dr) > T(0)) ? r : r / dr;
    out.radius = double(val(r));
    out.derivative = double(val(dr));
    out.iterations = done;
    out.escaped = escaped;
    out.z = V3<double>(double(val(z.x)), double(val(z.y)), double(val(z.z)));
    out.trap = trap;
    return out;
}

// Menger leaf (SM:756-784 with parametric count, size and divisor; no
// inter-level rotation).
template <class S, class T>
S menger_cut(V3<S> p, T s) {
    s *= T(2);
    auto md = [&](S x) {
        S y = x - S(s / T(2));
        return y - S(s) * S(std::floor(val(y) / s)) - S(s / T(2));
    };
    p = V3<S>(md(p.x), md(p.y), md(p.z));
    S da = s_max(s_abs(p.x), s_abs(p.y));
    S db = s_max(s_abs(p.y), s_abs(p.z));
    S dc = s_max(s_abs(p.z), s_abs(p.x));
    return s_min(da, s_min(db, dc)) - S(s / T(6));
}
template <class S, class T>
LeafResult<S> leaf_menger(V3<S> p, const float data[4]) {
    int iters = int(data[0]);
    T size = data[1] != 0.0f ? T(data[1]) : T(1);
    T divisor = data[2] != 0.0f ? T(data[2]) : T(3);
    V3<S> q = p / size;
    V3<S> d = vabs(q) - V3<S>(S(1), S(1), S(1));
    S sdf = s_max(s_max(d.x, d.y), d.z);
    T scale = T(1);
    for (int i = 0; i < iters; ++i) {
        sdf = s_max(sdf, -menger_cut<S, T>(q, scale));
        scale /= divisor;
    }
    count_orbit(uint64_t(iters));
    LeafResult<S> out;
    out.d = sdf * S(size);
    out.iterations = iters;
    out.trap = 0.0;
    return out;
}

template <class S, class T>
bool eval_leaf(uint32_t opcode, V3<S> p, const float data[4], LeafResult<S> &out) {
    switch (opcode) {
        case FPT_SDF_OP_MANDELBULB: out = leaf_mandelbulb<S, T>(p, data); return true;
        case FPT_SDF_OP_MANDELBOX: out = leaf_mandelbox<S, T>(p, data); return true;
        case FPT_SDF_OP_MENGER: out = leaf_menger<S, T>(p, data); return true;
        default: return false;
    }
}

// ---------------------------------------------------------------------------
// Typed program interpreter (SM:1065-1069, 1882-2036, 3266-3369).
// ---------------------------------------------------------------------------
template <class T>
T combine_distance(T current, T candidate, uint32_t mode) {  // SM:1065-1069
    if (mode == 1u) return m_max(current, candidate);
    if (mode == 2u) return m_max(current, -candidate);
    return m_min(current, candidate);
}

template <class T>
void rot2(T &a, T &b, T angle) {  // SM:454-458
    T s = std::sin(angle), c = std::cos(angle);
    T na = c * a - s * b;
    T nb = s * a + c * b;
    a = na;
    b = nb;
}

template <class T>
void box_fold(V3<T> &p, T limit) {
    p.x = m_clamp(p.x, -limit, limit) * T(2) - p.x;
    p.y = m_clamp(p.y, -limit, limit) * T(2) - p.y;
    p.z = m_clamp(p.z, -limit, limit) * T(2) - p.z;
}

// Transform ops shared by the distance and shading interpreters.
template <class T>
bool apply_transform(const FptSdfInstruction &op, V3<T> &p, T &distance_scale) {
    const float *d = op.data;
    switch (op.opcode) {
        case FPT_SDF_OP_ABS: p = vabs(p); return true;
        case FPT_SDF_OP_TRANSLATE: p = p - load3<T>(d); return true;
        case FPT_SDF_OP_SCALE: {
            T s = std::fabs(T(d[0])) > T(1.0e-6) ? T(d[0]) : T(1);
            p = p * s;
            distance_scale *= std::fabs(s);
            return true;
        }
        case FPT_SDF_OP_ROTATE_X: rot2(p.y, p.z, T(d[0])); return true;
        case FPT_SDF_OP_ROTATE_Y: rot2(p.x, p.z, T(d[0])); return true;
        case FPT_SDF_OP_ROTATE_Z: rot2(p.x, p.y, T(d[0])); return true;
        case FPT_SDF_OP_REPEAT: {
            V3<T> per(m_max(std::fabs(T(d[0])), T(1.0e-5)), m_max(std::fabs(T(d[1])), T(1.0e-5)),
                      m_max(std::fabs(T(d[2])), T(1.0e-5)));
            p.x -= per.x * std::floor(p.x / per.x + T(0.5));
            p.y -= per.y * std::floor(p.y / per.y + T(0.5));
            p.z -= per.z * std::floor(p.z / per.z + T(0.5));
            return true;
        }
        case FPT_SDF_OP_SORT_DESC:
            if (p.x < p.z) std::swap(p.x, p.z);
            if (p.y < p.z) std::swap(p.y, p.z);
            if (p.x < p.y) std::swap(p.x, p.y);
            return true;
        case FPT_SDF_OP_BOX_FOLD: {
            T l = d[0] != 0.0f ? T(d[0]) : T(1);
            box_fold(p, l);
            return true;
        }
        default: return false;
    }
}

// Primitive leaves sphere/box/plane (SM:2009-2031).
template <class T>
bool primitive_distance(const FptSdfInstruction &op, V3<T> p, T &out) {
    const float *d = op.data;
    switch (op.opcode) {
        case FPT_SDF_OP_SPHERE: out = length(p) - T(d[0]); return true;
        case FPT_SDF_OP_BOX: {
            V3<T> q = vabs(p) - vabs(load3<T>(d));
            V3<T> o(m_max(q.x, T(0)), m_max(q.y, T(0)), m_max(q.z, T(0)));
            out = m_min(m_max(q.x, m_max(q.y, q.z)), T(0)) + length(o);
            return true;
        }
        case FPT_SDF_OP_PLANE:
            out = dot(p, normalize(load3<T>(d))) + T(d[3]);
            return true;
        default: return false;
    }
}

struct ProgramSpan {
    const FptSdfInstruction *ops;
    uint32_t count;
};
ProgramSpan geometry_program(const FptRenderConfig &cfg) {
    return {cfg.sdf_program, std::min<uint32_t>(cfg.sdf_program_count, FPT_SDF_PROGRAM_MAX_OPS)};
}
ProgramSpan shading_program(const FptRenderConfig &cfg) {
    return {cfg.sdf_shading_program, std::min<uint32_t>(cfg.sdf_shading_program_count, FPT_SDF_PROGRAM_MAX_OPS)};
}

// deProgramDistance (SM:1961-2036) / deShadingProgram (SM:1882-1959).
template <class T>
T program_eval(ProgramSpan prog, V3<T> p, T *orbit_out) {
    T distance_scale = T(1);
    T distance = T(1.0e20);
    T orbit = T(0);
    bool has = false;
    for (uint32_t i = 0; i < prog.count; ++i) {
        const FptSdfInstruction &op = prog.ops[i];
        if (apply_transform(op, p, distance_scale)) continue;
        T cand;
        T div = m_max(distance_scale, T(1.0e-6));
        if (primitive_distance(op, p, cand)) {
            cand = cand / div;
            distance = has ? combine_distance(distance, cand, op.flags) : cand;
            has = true;
            if (op.opcode != FPT_SDF_OP_PLANE) orbit += std::fabs(cand) * T(op.data[3]);
            continue;
        }
        LeafResult<T> leaf;
        if (eval_leaf<T, T>(op.opcode, p, op.data, leaf)) {
            cand = leaf.d / div;
            distance = has ? combine_distance(distance, cand, op.flags) : cand;
            has = true;
            orbit += T(leaf.trap);
            continue;
        }
        if (op.opcode == FPT_SDF_OP_ORBIT_ADD) {
            orbit += length(p * load3<T>(op.data)) * T(op.data[3]);
        }
    }
    if (orbit_out) *orbit_out = orbit;
    return distance;
}

// programSurfaceInterpreted (SM:3266-3369), extended through fractal leaves
// by forward-mode differentiation of the leaf in its local frame.
template <class T>
T program_surface(ProgramSpan prog, V3<T> p, V3<T> &grad) {
    V3<T> jx(T(1), T(0), T(0)), jy(T(0), T(1), T(0)), jz(T(0), T(0), T(1));
    T distance_scale = T(1);
    T dist = T(1.0e20);
    grad = V3<T>();
    bool has = false;
    for (uint32_t i = 0; i < prog.count; ++i) {
        const FptSdfInstruction &op = prog.ops[i];
        const float *d = op.data;
        switch (op.opcode) {
            case FPT_SDF_OP_ABS:
                jx = jx * m_sign(p.x); jy = jy * m_sign(p.y); jz = jz * m_sign(p.z);
                p = vabs(p);
                continue;
            case FPT_SDF_OP_TRANSLATE: p = p - load3<T>(d); continue;
            case FPT_SDF_OP_SCALE: {
                T s = std::fabs(T(d[0])) > T(1.0e-6) ? T(d[0]) : T(1);
                p = p * s; jx = jx * s; jy = jy * s; jz = jz * s;
                distance_scale *= std::fabs(s);
                continue;
            }
            case FPT_SDF_OP_ROTATE_X: {
                T s = std::sin(T(d[0])), c = std::cos(T(d[0]));
                T py = p.y; V3<T> oj = jy;
                p.y = c * py - s * p.z; p.z = s * py + c * p.z;
                jy = oj * c - jz * s; jz = oj * s + jz * c;
                continue;
            }
            case FPT_SDF_OP_ROTATE_Y: {
                T s = std::sin(T(d[0])), c = std::cos(T(d[0]));
                T px = p.x; V3<T> oj = jx;
                p.x = c * px - s * p.z; p.z = s * px + c * p.z;
                jx = oj * c - jz * s; jz = oj * s + jz * c;
                continue;
            }
            case FPT_SDF_OP_ROTATE_Z: {
                T s = std::sin(T(d[0])), c = std::cos(T(d[0]));
                T px = p.x; V3<T> oj = jx;
                p.x = c * px - s * p.y; p.y = s * px + c * p.y;
                jx = oj * c - jy * s; jy = oj * s + jy * c;
                continue;
            }
            case FPT_SDF_OP_REPEAT: {
                T dist_scale_unused = T(1);
                apply_transform(op, p, dist_scale_unused);
                continue;
            }
            case FPT_SDF_OP_SORT_DESC:
                if (p.x < p.z) { std::swap(p.x, p.z); std::swap(jx, jz); }
                if (p.y < p.z) { std::swap(p.y, p.z); std::swap(jy, jz); }
                if (p.x < p.y) { std::swap(p.x, p.y); std::swap(jx, jy); }
                continue;
            case FPT_SDF_OP_BOX_FOLD: {
                // d/dp of clamp(p,-l,l)*2 - p: +1 inside the fold box, -1 outside.
                T l = d[0] != 0.0f ? T(d[0]) : T(1);
                T fx = std::fabs(p.x) > l ? T(-1) : T(1);
                T fy = std::fabs(p.y) > l ? T(-1) : T(1);
                T fz = std::fabs(p.z) > l ? T(-1) : T(1);
                box_fold(p, l);
                jx = jx * fx; jy = jy * fy; jz = jz * fz;
                continue;
            }
            default: break;
        }
        T scale = m_max(distance_scale, T(1.0e-6));
        T cand;
        V3<T> lg;
        if (op.opcode == FPT_SDF_OP_SPHERE) {
            T radius = length(p);
            cand = (radius - T(d[0])) / scale;
            lg = radius > T(1.0e-8) ? p / radius : V3<T>(T(0), T(1), T(0));
        } else if (op.opcode == FPT_SDF_OP_BOX) {
            V3<T> q = vabs(p) - vabs(load3<T>(d));
            V3<T> o(m_max(q.x, T(0)), m_max(q.y, T(0)), m_max(q.z, T(0)));
            T ol = length(o);
            cand = (m_min(m_max(q.x, m_max(q.y, q.z)), T(0)) + ol) / scale;
            if (ol > T(1.0e-8)) {
                lg = V3<T>(m_sign(p.x), m_sign(p.y), m_sign(p.z)) * o / ol;
            } else if (q.x >= q.y && q.x >= q.z) {
                lg = V3<T>(m_sign(p.x), T(0), T(0));
            } else if (q.y >= q.z) {
                lg = V3<T>(T(0), m_sign(p.y), T(0));
            } else {
                lg = V3<T>(T(0), T(0), m_sign(p.z));
            }
        } else if (op.opcode == FPT_SDF_OP_PLANE) {
            lg = normalize(load3<T>(d));
            cand = (dot(p, lg) + T(d[3])) / scale;
        } else if (op.opcode == FPT_SDF_OP_MANDELBULB || op.opcode == FPT_SDF_OP_MANDELBOX ||
                   op.opcode == FPT_SDF_OP_MENGER) {
            V3<Dual<T>> dp(Dual<T>(p.x, T(1), T(0), T(0)), Dual<T>(p.y, T(0), T(1), T(0)),
                           Dual<T>(p.z, T(0), T(0), T(1)));
            LeafResult<Dual<T>> leaf;
            eval_leaf<Dual<T>, T>(op.opcode, dp, d, leaf);
            cand = leaf.d.v / scale;
            lg = V3<T>(leaf.d.d0, leaf.d.d1, leaf.d.d2);
        } else {
            continue;
        }
        V3<T> cg = (jx * lg.x + jy * lg.y + jz * lg.z) / scale;
        bool choose = !has || (op.flags == 1u ? cand > dist
                                              : (op.flags == 2u ? -cand > dist : cand < dist));
        if (choose) {
            dist = (op.flags == 2u && has) ? -cand : cand;
            grad = (op.flags == 2u && has) ? -cg : cg;
        }
        has = true;
    }
    return dist;
}

// ---------------------------------------------------------------------------
// Built-in fractal scenes used by the configs (SM:730-784).
// ---------------------------------------------------------------------------
template <class T>
T de_mandelbox_builtin(V3<T> pos, const FptRenderConfig &cfg, T *orbit) {  // SM:730-754
    T SCALE = T(2.8f) + T(cfg.set_values[2]);
    T MINRAD2 = T(0.25f) + T(cfg.set_values[3]);
    T sc[4] = {SCALE / MINRAD2, SCALE / MINRAD2, SCALE / MINRAD2, std::fabs(SCALE) / MINRAD2};
    T min_rad2 = m_clamp(MINRAD2, T(1.0e-9), T(1));
    T abs_scale_m1 = std::fabs(SCALE - T(1));
    // |SCALE|^-9 as the IEEE chain ((s^2)^2)^2 * s (the GPU uses the same).
    T as1 = std::fabs(SCALE), as2 = as1 * as1, as4 = as2 * as2, as8 = as4 * as4;
    T abs_scale = T(1) / (as8 * as1);
    T off[4] = {T(cfg.set_values[4]), T(cfg.set_values[5]), T(cfg.set_values[6]), T(0)};
    T p[4] = {pos.x, pos.y, pos.z, T(1)};
    T p0[4] = {p[0], p[1], p[2], p[3]};
    T sdf = 0, col = 0;
    for (int i = 0; i < 12; ++i) {
        for (int a = 0; a < 3; ++a) p[a] = m_clamp(p[a], T(-1), T(1)) * T(2) - p[a];
        T r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        T k = m_clamp(m_max(min_rad2 / m_max(r2, T(0.00001)), min_rad2), T(0), T(1));
        for (int a = 0; a < 4; ++a) p[a] *= k;
        for (int a = 0; a < 4; ++a) p[a] = p[a] * sc[a] + p0[a] + off[a];
        T len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        sdf = ((len - abs_scale_m1) / p[3] - abs_scale);
        col += sdf;
    }
    count_orbit(12);
    if (orbit) *orbit = col;
    return sdf;
}

template <class T>
T de_menger_builtin(V3<T> p, const FptRenderConfig &cfg) {  // SM:756-784
    V3<T> dd = vabs(p) - V3<T>(T(1), T(1), T(1));
    T sdf = m_max(m_max(dd.x, dd.y), dd.z);
    T scale = m_max(T(cfg.set_values[0]), T(0.0001));
    T divisor = m_max(T(cfg.set_values[1]), T(0.0001));
    T ry = T(cfg.set_values[2]), rp = T(cfg.set_values[3]);
    for (int i = 0; i < 12; ++i) {
        T s = scale * T(2);
        auto md = [&](T x) { T y = x - s / T(2); return y - s * std::floor(y / s) - s / T(2); };
        V3<T> q(md(p.x), md(p.y), md(p.z));
        T da = m_max(std::fabs(q.x), std::fabs(q.y));
        T db = m_max(std::fabs(q.y), std::fabs(q.z));
        T dc = m_max(std::fabs(q.z), std::fabs(q.x));
        T cut = m_min(da, m_min(db, dc)) - s / T(6);
        sdf = m_max(sdf, -cut);
        scale /= divisor;
        p = rotate_camera(p, ry, rp);
    }
    count_orbit(12);
    return sdf;
}

// ---------------------------------------------------------------------------
// Materials and scene dispatch (SM:3418-3615).
// ---------------------------------------------------------------------------
template <class T>
struct Material {
    V3<T> rgb{T(1), T(1), T(1)};
    T roughness = T(1), specular = T(0), translucency = T(0), ior = T(1.5), emission = T(0);
};

template <class T>
V3<T> hsv2rgb(V3<T> c) {  // SM:534-539
    T K[4] = {T(1), T(2) / T(3), T(1) / T(3), T(3)};
    V3<T> p(std::fabs(m_fract(c.x + K[0]) * T(6) - K[3]), std::fabs(m_fract(c.x + K[1]) * T(6) - K[3]),
            std::fabs(m_fract(c.x + K[2]) * T(6) - K[3]));
    V3<T> q(m_clamp(p.x - K[0], T(0), T(1)), m_clamp(p.y - K[0], T(0), T(1)), m_clamp(p.z - K[0], T(0), T(1)));
    return vmix(V3<T>(K[0], K[0], K[0]), q, c.y) * c.z;
}

template <class T>
V3<T> program_gradient(T value, const FptRenderConfig &cfg) {  // SM:3418-3432
    uint32_t count = std::min<uint32_t>(cfg.gradient_count, FPT_SDF_GRADIENT_MAX_STOPS);
    if (count == 0) return load3<T>(cfg.program_material);
    if (count == 1) return load3<T>(&cfg.gradient_stops[0][1]);
    T t = m_fract(value);
    const float *prev = cfg.gradient_stops[0];
    for (uint32_t i = 1; i < count; ++i) {
        const float *cur = cfg.gradient_stops[i];
        if (t <= T(cur[0])) {
            T f = m_clamp((t - T(prev[0])) / m_max(T(cur[0]) - T(prev[0]), T(1.0e-6)), T(0), T(1));
            return vmix(load3<T>(prev + 1), load3<T>(cur + 1), f);
        }
        prev = cur;
    }
    return load3<T>(prev + 1);
}

template <class T>
bool is_program(const FptRenderConfig &cfg) {
    return cfg.sdf_id == FPT_SDF_PROGRAM && cfg.sdf_program_count > 0;
}

#include "fpt_oracle_kifs.inc"

// Mandelbulber march threshold (SM:3634-3640): fixed (vset_values[117]) or
// scaled by the camera distance (mode [115] > 0.5, factor [116]), clamped
// to [118..119].
template <class T>
T mb_threshold(V3<T> p, const FptRenderConfig &cfg) {
    const float *v = cfg.vset_values;
    const T t = v[115] > 0.5f ? length(load3<T>(cfg.camera_position) - p) * T(v[116]) : T(v[117]);
    return m_clamp(t, T(v[118]), T(v[119]));
}

// deMandelbulber (SM:975-995): in iteration-limit mode (vset_values[115] >
// 1.5) points that reach the iteration limit are interior (distance 0) and
// escaped points stay just outside the threshold; orbit = |iterations| /
// configured iterations.
template <class T>
T de_kifs(V3<T> p, const FptRenderConfig &cfg, T *orbit) {
    T s[4];
    kifs_field_sample<T>(p, cfg, 1, s);
    count_orbit(uint64_t(std::fabs(s[3])));
    T d = s[0];
    const float *v = cfg.vset_values;
    if (v[115] > 1.5f) {
        const T thr = m_clamp(length(load3<T>(cfg.camera_position) - p) * T(v[116]), T(v[118]), T(v[119]));
        if (s[3] > T(0)) d = T(0);
        else if (d < thr) d = thr * T(1.01f);
    }
    if (orbit) *orbit = std::fabs(s[3]) / m_max(T(cfg.set_values[1]), T(1));
    return d;
}

// ---------------------------------------------------------------------------
// Direct evaluators of lowered programs (SM:2713-2970): typed
// structure-of-arrays, canonical affine and flat-union tables.
// ---------------------------------------------------------------------------
inline uint32_t lowered_cap(uint32_t n) { return std::min<uint32_t>(n, FPT_SDF_FLAT_UNION_MAX_PRIMITIVES); }
inline uint32_t typed_soa_count(const FptRenderConfig &cfg) {  // SM:2848-2855
    const FptTypedSoAProgram &t = cfg.sdf_typed_soa;
    return lowered_cap(t.sphere_count) + lowered_cap(t.box_count) + lowered_cap(t.plane_count);
}

template <class T>
V3<T> instance_point(V3<T> s, const float *t) {
    return V3<T>(dot(V3<T>(T(t[0]), T(t[1]), T(t[2])), s) + T(t[3]),
                 dot(V3<T>(T(t[4]), T(t[5]), T(t[6])), s) + T(t[7]),
                 dot(V3<T>(T(t[8]), T(t[9]), T(t[10])), s) + T(t[11]));
}

template <class T>
T instance_distance(V3<T> s, const FptPrimitiveInstance &in) {  // SM:2713-2737
    const V3<T> p = instance_point(s, in.transform);
    const T div = m_max(T(in.distance_scale), T(1.0e-6f));
    const V3<T> data(T(in.data[0]), T(in.data[1]), T(in.data[2]));
    if (in.opcode == FPT_SDF_OP_SPHERE) return (length(p) - data.x) / div;
    if (in.opcode == FPT_SDF_OP_BOX) {
        const V3<T> q = vabs(p) - vabs(data);
        const V3<T> o(m_max(q.x, T(0)), m_max(q.y, T(0)), m_max(q.z, T(0)));
        return (m_min(m_max(q.x, m_max(q.y, q.z)), T(0)) + length(o)) / div;
    }
    if (in.opcode == FPT_SDF_OP_PLANE) return (dot(p, data) + T(in.data[3])) / div;
    return T(1.0e20);
}

template <class T>
T flat_union_distance(V3<T> s, const FptRenderConfig &cfg) {  // SM:2739-2749
    T d = T(1.0e20);
    for (uint32_t i = 0, n = lowered_cap(cfg.sdf_flat_union_count); i < n; ++i)
        d = m_min(d, instance_distance(s, cfg.sdf_flat_union_instances[i]));
    return d;
}

template <class T>
T canonical_distance(V3<T> s, const FptRenderConfig &cfg) {  // SM:2751-2770: _pad0 0 union, 1 and, 2 minus
    T d = T(1.0e20);
    for (uint32_t i = 0, n = lowered_cap(cfg.sdf_canonical_count); i < n; ++i) {
        const FptPrimitiveInstance &in = cfg.sdf_canonical_primitives[i];
        const T c = instance_distance(s, in);
        if (i == 0u) d = c;
        else if (in._pad0 == 1u) d = m_max(d, c);
        else if (in._pad0 == 2u) d = m_max(d, -c);
        else d = m_min(d, c);
    }
    return d;
}

template <class T>
T canonical_surface(V3<T> s, const FptRenderConfig &cfg, V3<T> &grad) {  // SM:2772-2846
    T d = T(1.0e20);
    grad = V3<T>();
    for (uint32_t i = 0, n = lowered_cap(cfg.sdf_canonical_count); i < n; ++i) {
        const FptPrimitiveInstance &in = cfg.sdf_canonical_primitives[i];
        const float *t = in.transform;
        const V3<T> r0 = V3<T>(T(t[0]), T(t[1]), T(t[2]));
        const V3<T> r1 = V3<T>(T(t[4]), T(t[5]), T(t[6]));
        const V3<T> r2 = V3<T>(T(t[8]), T(t[9]), T(t[10]));
        const V3<T> p(dot(r0, s) + T(t[3]), dot(r1, s) + T(t[7]), dot(r2, s) + T(t[11]));
        const T div = m_max(T(in.distance_scale), T(1.0e-6f));
        T c = T(1.0e20);
        V3<T> lg(T(0), T(1), T(0));
        if (in.opcode == FPT_SDF_OP_SPHERE) {
            const T r = length(p);
            c = (r - T(in.data[0])) / div;
            if (r > T(1.0e-8f)) lg = p / r;
        } else if (in.opcode == FPT_SDF_OP_BOX) {
            const V3<T> q = vabs(p) - vabs(V3<T>(T(in.data[0]), T(in.data[1]), T(in.data[2])));
            const V3<T> o(m_max(q.x, T(0)), m_max(q.y, T(0)), m_max(q.z, T(0)));
            const T ol = length(o);
            c = (m_min(m_max(q.x, m_max(q.y, q.z)), T(0)) + ol) / div;
            if (ol > T(1.0e-8f)) lg = V3<T>(m_sign(p.x), m_sign(p.y), m_sign(p.z)) * o / ol;
            else if (q.x >= q.y && q.x >= q.z) lg = V3<T>(m_sign(p.x), T(0), T(0));
            else if (q.y >= q.z) lg = V3<T>(T(0), m_sign(p.y), T(0));
            else lg = V3<T>(T(0), T(0), m_sign(p.z));
        } else if (in.opcode == FPT_SDF_OP_PLANE) {
            lg = V3<T>(T(in.data[0]), T(in.data[1]), T(in.data[2]));
            c = (dot(p, lg) + T(in.data[3])) / div;
        }
        const V3<T> cg = (r0 * lg.x + r1 * lg.y + r2 * lg.z) / div;
        const bool choose = i == 0u || (in._pad0 == 1u ? c > d : (in._pad0 == 2u ? -c > d : c < d));
        if (choose) {
            const bool neg = in._pad0 == 2u && i > 0u;
            d = neg ? -c : c;
            grad = neg ? -cg : cg;
        }
    }
    return d;
}

template <class T>
T typed_soa_distance(V3<T> s, const FptRenderConfig &cfg) {  // SM:2857-2896
    const FptTypedSoAProgram &t = cfg.sdf_typed_soa;
    T d = T(1.0e20);
    for (uint32_t i = 0, n = lowered_cap(t.sphere_count); i < n; ++i)
        d = m_min(d, length(s - V3<T>(T(t.sphere_x[i]), T(t.sphere_y[i]), T(t.sphere_z[i]))) - T(t.sphere_radius[i]));
    for (uint32_t i = 0, n = lowered_cap(t.box_count); i < n; ++i) {
        const V3<T> q = vabs(s - V3<T>(T(t.box_x[i]), T(t.box_y[i]), T(t.box_z[i]))) -
                        V3<T>(T(t.box_half_x[i]), T(t.box_half_y[i]), T(t.box_half_z[i]));
        const V3<T> o(m_max(q.x, T(0)), m_max(q.y, T(0)), m_max(q.z, T(0)));
        d = m_min(d, m_min(m_max(q.x, m_max(q.y, q.z)), T(0)) + length(o));
    }
    for (uint32_t i = 0, n = lowered_cap(t.plane_count); i < n; ++i) {
        const V3<T> nrm = normalize(V3<T>(T(t.plane_x[i]), T(t.plane_y[i]), T(t.plane_z[i])));
        const V3<T> c(T(t.plane_center_x[i]), T(t.plane_center_y[i]), T(t.plane_center_z[i]));
        d = m_min(d, dot(s - c, nrm) + T(t.plane_offset[i]));
    }
    return d;
}

template <class T>
T typed_soa_surface(V3<T> s, const FptRenderConfig &cfg, V3<T> &grad) {  // SM:2898-2970: ties to the earlier source
    const FptTypedSoAProgram &t = cfg.sdf_typed_soa;
    T d = T(1.0e20);
    grad = V3<T>();
    uint32_t winner = 0xffffffffu;
    for (uint32_t i = 0, n = lowered_cap(t.sphere_count); i < n; ++i) {
        const V3<T> l = s - V3<T>(T(t.sphere_x[i]), T(t.sphere_y[i]), T(t.sphere_z[i]));
        const T r = length(l), c = r - T(t.sphere_radius[i]);
        if (c < d || (c == d && t.sphere_source[i] < winner)) {
            d = c;
            grad = r > T(1.0e-8f) ? l / r : V3<T>(T(0), T(1), T(0));
            winner = t.sphere_source[i];
        }
    }
    for (uint32_t i = 0, n = lowered_cap(t.box_count); i < n; ++i) {
        const V3<T> l = s - V3<T>(T(t.box_x[i]), T(t.box_y[i]), T(t.box_z[i]));
        const V3<T> q = vabs(l) - V3<T>(T(t.box_half_x[i]), T(t.box_half_y[i]), T(t.box_half_z[i]));
        const V3<T> o(m_max(q.x, T(0)), m_max(q.y, T(0)), m_max(q.z, T(0)));
        const T ol = length(o);
        const T c = m_min(m_max(q.x, m_max(q.y, q.z)), T(0)) + ol;
        if (c < d || (c == d && t.box_source[i] < winner)) {
            d = c;
            if (ol > T(1.0e-8f)) grad = V3<T>(m_sign(l.x), m_sign(l.y), m_sign(l.z)) * o / ol;
            else if (q.x >= q.y && q.x >= q.z) grad = V3<T>(m_sign(l.x), T(0), T(0));
            else if (q.y >= q.z) grad = V3<T>(T(0), m_sign(l.y), T(0));
            else grad = V3<T>(T(0), T(0), m_sign(l.z));
            winner = t.box_source[i];
        }
    }
    for (uint32_t i = 0, n = lowered_cap(t.plane_count); i < n; ++i) {
        const V3<T> nrm = normalize(V3<T>(T(t.plane_x[i]), T(t.plane_y[i]), T(t.plane_z[i])));
        const V3<T> c0(T(t.plane_center_x[i]), T(t.plane_center_y[i]), T(t.plane_center_z[i]));
        const T c = dot(s - c0, nrm) + T(t.plane_offset[i]);
        if (c < d || (c == d && t.plane_source[i] < winner)) {
            d = c;
            grad = nrm;
            winner = t.plane_source[i];
        }
    }
    return d;
}

template <class T>
T distance_sdf(V3<T> p, const FptRenderConfig &cfg) {  // SM:3564-3610
    if (tl_counters) tl_counters->distance_evals++;
    if (is_program<T>(cfg)) {  // lowered tables first (SM:3576-3584)
        if (typed_soa_count(cfg) > 0u) return typed_soa_distance(p, cfg);
        if (cfg.sdf_canonical_count > 0u) return canonical_distance(p, cfg);
        if (cfg.sdf_flat_union_count > 0u) return flat_union_distance(p, cfg);
    }
    if (is_program<T>(cfg)) return program_eval<T>(geometry_program(cfg), p, nullptr);
    switch (cfg.sdf_id) {
        case FPT_SDF_MANDELBULBER: return de_kifs<T>(p, cfg, nullptr);
        case FPT_SDF_MANDELBOX_FRACTAL: return de_mandelbox_builtin<T>(p, cfg, nullptr);
        case FPT_SDF_MENGER_SPONGE: return de_menger_builtin<T>(p, cfg);
        default: return T(1000);
    }
}

template <class T>
T map_sdf(V3<T> p, const FptRenderConfig &cfg) {  // SM:3612-3615
    T d = distance_sdf(p, cfg);
    return m_mix(d, d / T(2), T(cfg.render[6]));
}

// Material as orbit -> shading parameter -> material (SM:3434-3483), the
// split the voxel palette needs; user_material composes the three.
template <class T>
T material_orbit(V3<T> p, const FptRenderConfig &cfg) {  // SM:3485-3562 (orbit part)
    T orbit = T(0);
    if (is_program<T>(cfg)) {
        if (cfg.material_mode == 1u || cfg.material_mode == 2u) program_eval<T>(shading_program(cfg), p, &orbit);
        return orbit;
    }
    if (cfg.sdf_id == FPT_SDF_MANDELBOX_FRACTAL) de_mandelbox_builtin<T>(p, cfg, &orbit);
    else if (cfg.sdf_id == FPT_SDF_MENGER_SPONGE) orbit = T(1);
    else if (cfg.sdf_id == FPT_SDF_MANDELBULBER && cfg.fractal_style_mode != 2u) de_kifs<T>(p, cfg, &orbit);
    return orbit;
}

template <class T>
T material_param(T orbit, const FptRenderConfig &cfg) {
    if (!is_program<T>(cfg) && cfg.fractal_style_mode != 0u)
        return orbit * T(cfg.fractal_style[1]) + T(cfg.fractal_style[0]);
    return orbit * (T(1) + T(cfg.set_values[1])) + T(cfg.set_values[0]);
}

template <class T>
bool material_is_constant(const FptRenderConfig &cfg) {
    if (is_program<T>(cfg)) return cfg.material_mode != 1u && cfg.material_mode != 2u;
    if (cfg.fractal_style_mode == 2u) return true;
    return cfg.fractal_style_mode == 0u && cfg.sdf_id == FPT_SDF_MENGER_SPONGE;
}

template <class T>
Material<T> material_at_param(T o, const FptRenderConfig &cfg) {
    Material<T> m;
    if (is_program<T>(cfg)) {  // SM:3434-3449
        if (cfg.material_mode == 1u) m.rgb = hsv2rgb(V3<T>(o, T(0.5), T(1)));
        else if (cfg.material_mode == 2u) m.rgb = program_gradient(o, cfg);
        else m.rgb = load3<T>(cfg.program_material);
        m.roughness = T(cfg.program_material[3]);
        m.specular = T(cfg.program_material[4]);
        m.translucency = T(cfg.program_material[5]);
        m.ior = T(cfg.program_material[6]);
        m.emission = T(cfg.program_material[7]);
        return m;
    }
    if (cfg.fractal_style_mode != 0u) {  // SM:3451-3468
        if (cfg.fractal_style_mode == 1u)
            m.rgb = hsv2rgb(V3<T>(o, T(cfg.fractal_style[2]), T(cfg.fractal_style[3])));
        else if (cfg.fractal_style_mode == 2u)
            m.rgb = load3<T>(&cfg.fractal_style[8]) * T(cfg.fractal_style[3]);
        else
            m.rgb = program_gradient(o, cfg) * T(cfg.fractal_style[3]);
        m.rgb = vmix(m.rgb, load3<T>(&cfg.fractal_style[8]), m_clamp(T(cfg.fractal_style[7]), T(0), T(1)));
        m.roughness = T(cfg.fractal_style[4]);
        m.specular = T(cfg.fractal_style[5]);
        m.emission = T(cfg.fractal_style[6]);
        return m;
    }
    if (cfg.sdf_id == FPT_SDF_MENGER_SPONGE) return m;
    m.rgb = hsv2rgb(V3<T>(o, T(0.5), T(1)));
    return m;
}

template <class T>
Material<T> user_material(V3<T> p, const FptRenderConfig &cfg) {  // SM:3485-3562 (material part)
    return material_at_param(material_param(material_orbit(p, cfg), cfg), cfg);
}

template <class T>
bool has_translucent(const FptRenderConfig &cfg) {  // SM:3617-3624 (program scenes)
    return cfg.sdf_id == FPT_SDF_PROGRAM && cfg.program_material[5] > 0.0f;
}

// marchMandelbulber (SM:3700-3751): march until d < threshold with the
// clamped Mandelbulber step (SM:3642-3676), stopping where the point no
// longer moves or max_distance is travelled, then up to 30 half-steps
// towards the band (threshold * (1 - 0.001 detail), threshold); steps counts
// the march iterations.
template <class T>
T mb_step(T d, T thr, const FptRenderConfig &cfg) {
    const float *v = cfg.vset_values;
    T step = m_max(d - T(0.5f) * thr, T(0)) * T(v[113]);
    if (v[108] > 0.5f) {
        step = m_clamp(step, T(v[109]), T(v[110]));
        if (thr > T(v[109])) step = m_clamp(step, T(v[111]) * thr, T(v[112]) * thr);
    } else {
        step = m_min(step, T(v[110]));
    }
    return step;
}

template <class T>
bool same_point(V3<T> a, V3<T> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

template <class T>
V3<T> march_mandelbulber(V3<T> dr, V3<T> rp, int ni, const FptRenderConfig &cfg, int *steps_out,
                         bool *found_out = nullptr) {
    const V3<T> start = rp;
    T d = T(0), thr = mb_threshold(rp, cfg), step = T(0);
    bool found = false;
    int steps = 0;
    const int max_iter = std::min(ni, 10000);
    for (int i = 0; i < max_iter; ++i) {
        thr = mb_threshold(rp, cfg);
        d = map_sdf(rp, cfg);
        ++steps;
        if (!std::isfinite(d)) break;
        if (d < thr) {
            found = true;
            break;
        }
        step = mb_step(d, thr, cfg);
        const V3<T> nx = rp + dr * step;
        if (same_point(nx, rp)) break;
        rp = nx;
        if (length(rp - start) >= T(cfg.render[4])) break;
    }
    if (steps_out) *steps_out = steps;
    if (found_out) *found_out = found;
    if (!found) return rp;
    const T limit = T(1) - T(0.001f) * m_max(T(cfg.vset_values[129]), T(0));
    step *= T(0.5f);
    for (int r = 0; r < 30; ++r) {
        if (d < thr && d > thr * limit) break;
        if (d > thr) {
            const V3<T> nx = rp + dr * step;
            if (same_point(nx, rp)) break;
            rp = nx;
        } else if (d < thr * limit) {
            const V3<T> nx = rp - dr * step;
            if (same_point(nx, rp)) break;
            rp = nx;
        }
        d = map_sdf(rp, cfg);
        step *= T(0.5f);
    }
    return rp;
}

// march (SM:3753-3783); Mandelbulber scenes take marchMandelbulber.
template <class T>
V3<T> march(V3<T> dr, V3<T> rp, int ni, T min_dist, T lod_falloff, const FptRenderConfig &cfg,
            int *steps_out) {
    if (cfg.sdf_id == FPT_SDF_MANDELBULBER) return march_mandelbulber(dr, rp, ni, cfg, steps_out);
    V3<T> cam = rp;
    int max_iter = std::min(ni, 360);
    bool check_t = has_translucent<T>(cfg);
    int steps = 0;
    for (int i = 0; i < max_iter; ++i) {
        uint64_t it0 = tl_counters ? tl_counters->orbit_iterations : 0;
        T d = map_sdf(rp, cfg);
        ++steps;
        if (tl_counters && std::fabs(d) > T(0.02)) {
            tl_counters->far_steps++;
            tl_counters->far_orbit_iterations += tl_counters->orbit_iterations - it0;
        }
        if (!std::isfinite(d)) break;
        T thr = min_dist;
        T o = std::fabs(d) * T(0.99f);
        rp = rp + dr * o;
        T lod = thr;
        if (lod_falloff > T(0.00001f)) {
            V3<T> df = cam - rp;
            T fog = dot(df, df);
            lod = m_mix(lod, T(0.1f), fog * lod_falloff);
            lod = m_mix(thr, lod, T(cfg.render[5]));
        }
        if (check_t && user_material(rp, cfg).translucency > T(0)) lod = T(0.0001f);
        if (o < lod) break;
        if (T(cfg.render[4]) < o) break;
    }
    if (steps_out) *steps_out = steps;
    return rp;
}

// estimateFocusDistance (SM:5869-5878): hit distance of the un-jittered
// centre ray (120 steps, threshold and LOD falloff 0.001), or 5 when it is
// not finite, within 0.001 of the camera or beyond 0.99 max_distance.
template <class T>
T focus_estimate(const FptRenderConfig &cfg) {
    const V3<T> dr = camera_dir<T>(T(0), T(0), cfg);
    const V3<T> cam = load3<T>(cfg.camera_position);
    const V3<T> hit = march(dr, cam, 120, T(0.001f), T(0.001f), cfg, nullptr);
    const T d = length(hit - cam);
    if (!std::isfinite(d) || d <= T(0.001f) || d > T(cfg.render[4]) * T(0.99f)) return T(5);
    return d;
}

// Thin-lens camera (SM:5950-5956): keep the ray's point at the focus
// distance, start it from a lens offset of radius camera_dof.
template <class T>
void apply_dof(UV xy, float frame, const FptRenderConfig &cfg, V3<T> &rp, V3<T> &dr) {
    if (!(cfg.camera_dof > 0.0f)) return;
    const T focus = cfg.focus_distance > 0.0f ? T(cfg.focus_distance) : focus_estimate<T>(cfg);
    const V3<T> fp = rp + dr * focus;
    T lx, ly;
    random_point<T>(T(cfg.camera_dof), xy, frame, lx, ly);
    rp = rp + rotate_camera_roll(V3<T>(lx, ly, T(0)), T(cfg.camera_yaw_pitch[0]), T(cfg.camera_yaw_pitch[1]),
                                 T(cfg.camera_roll));
    dr = normalize(fp - rp);
}

// The render entry points' focus prepass (metal_bridge.mm:4827-4838,
// 5455-5481): a config with depth of field and no focus distance gets the
// estimate, rounded to the config's float field.
template <class T>
FptRenderConfig with_focus(const FptRenderConfig &cfg) {
    FptRenderConfig c = cfg;
    if (c.camera_dof > 0.0f && !(c.focus_distance > 0.0f)) c.focus_distance = float(focus_estimate<T>(c));
    return c;
}

template <class T>
V3<T> normalize_fd(V3<T> g) {  // SM:3863-3869
    T m = m_max(std::fabs(g.x), m_max(std::fabs(g.y), std::fabs(g.z)));
    if (!(m > T(0)) || !std::isfinite(m)) return V3<T>(T(0), T(1), T(0));
    return normalize(g / m);
}

// The Mandelbulber normal stencil's field sample (SM:964-974): 5x the
// iterations in iteration-limit mode, the raw field distance.
template <class T>
void mb_normal_sample(V3<T> p, const FptRenderConfig &cfg, int mult, T s[4]) {
    if (tl_counters) tl_counters->distance_evals++;
    kifs_field_sample<T>(p, cfg, mult, s);
    count_orbit(uint64_t(std::fabs(s[3])));
}

// normalAt's Mandelbulber branches (SM:3871-3960): the slow-shading 11^3
// iteration-count stencil when render[7] > 0; otherwise the tetrahedral or
// central stencil at threshold * vset_values[114], with the log orbit
// radius as the fallback field when the six central samples are equal.
template <class T>
V3<T> normal_mandelbulber(V3<T> p, const FptRenderConfig &cfg) {
    const float *v = cfg.vset_values;
    T s[4];
    if (cfg.render[7] > 0.0f) {
        const T delta = length(load3<T>(cfg.camera_position) - p) * T(cfg.render[7]) * T(v[114]);
        const T max_it = T(cfg.set_values[1]);
        V3<T> g;
        for (int ix = -5; ix <= 5; ++ix)
            for (int iy = -5; iy <= 5; ++iy)
                for (int iz = -5; iz <= 5; ++iz) {
                    const V3<T> sd = V3<T>(T(ix), T(iy), T(iz)) * T(0.2f);
                    mb_normal_sample<T>(p + sd * delta, cfg, 1, s);
                    g = g + sd * (T(1) + max_it - std::fabs(s[3]));
                }
        const T m = m_max(std::fabs(g.x), m_max(std::fabs(g.y), std::fabs(g.z)));
        if (!(m > T(0)) || !std::isfinite(m)) return V3<T>(T(1), T(0), T(0));
        return normalize(g / m);
    }
    const int mult = v[115] > 1.5f ? 5 : 1;
    auto dist = [&](V3<T> q) { mb_normal_sample<T>(q, cfg, mult, s); return s[0]; };
    const T e = mb_threshold(p, cfg) * T(v[114]);
    if (cfg.sdf_normal_mode == 1u) {
        const T te = e * T(0.57735026919f);
        const V3<T> k0(T(1), T(-1), T(-1)), k1(T(-1), T(-1), T(1)), k2(T(-1), T(1), T(-1)), k3(T(1), T(1), T(1));
        const T d0 = dist(p + k0 * te), d1 = dist(p + k1 * te), d2 = dist(p + k2 * te), d3 = dist(p + k3 * te);
        return normalize_fd(k0 * d0 + k1 * d1 + k2 * d2 + k3 * d3);
    }
    const V3<T> ex(e, T(0), T(0)), ey(T(0), e, T(0)), ez(T(0), T(0), e);
    V3<T> n(dist(p + ex) - dist(p - ex), dist(p + ey) - dist(p - ey), dist(p + ez) - dist(p - ez));
    if (!(m_max(std::fabs(n.x), m_max(std::fabs(n.y), std::fabs(n.z))) > T(0))) {
        auto lr = [&](V3<T> q) {
            mb_normal_sample<T>(q, cfg, mult, s);
            return std::log(m_max(std::fabs(s[1]), T(1.0e-30)));
        };
        n = V3<T>(lr(p + ex) - lr(p - ex), lr(p + ey) - lr(p - ey), lr(p + ez) - lr(p - ez));
    }
    return normalize_fd(n);
}

// normalAt (SM:3871-3971).
template <class T>
V3<T> normal_at(V3<T> p, const FptRenderConfig &cfg) {
    if (tl_counters) tl_counters->normal_evals++;
    if (cfg.sdf_id == FPT_SDF_MANDELBULBER) return normal_mandelbulber(p, cfg);
    T e = m_max(T(cfg.render[2]), T(0.0002f));
    if ((cfg.sdf_normal_mode == 0u || cfg.sdf_normal_mode == 2u) && is_program<T>(cfg)) {
        V3<T> g;  // programSurface (SM:3403-3408): typed SoA, canonical, else interpreted
        if (typed_soa_count(cfg) > 0u) typed_soa_surface<T>(p, cfg, g);
        else if (cfg.sdf_canonical_count > 0u) canonical_surface<T>(p, cfg, g);
        else program_surface<T>(geometry_program(cfg), p, g);
        if (dot(g, g) >= T(1.0e-12f) && vfinite(g)) return normalize(g);
    }
    if (cfg.sdf_normal_mode == 1u) {
        T te = e * T(0.57735026919f);
        V3<T> k0(T(1), T(-1), T(-1)), k1(T(-1), T(-1), T(1)), k2(T(-1), T(1), T(-1)), k3(T(1), T(1), T(1));
        T d0 = map_sdf(p + k0 * te, cfg), d1 = map_sdf(p + k1 * te, cfg);
        T d2 = map_sdf(p + k2 * te, cfg), d3 = map_sdf(p + k3 * te, cfg);
        return normalize_fd(k0 * d0 + k1 * d1 + k2 * d2 + k3 * d3);
    }
    V3<T> ex(e, T(0), T(0)), ey(T(0), e, T(0)), ez(T(0), T(0), e);
    V3<T> n(map_sdf(p + ex, cfg) - map_sdf(p - ex, cfg), map_sdf(p + ey, cfg) - map_sdf(p - ey, cfg),
            map_sdf(p + ez, cfg) - map_sdf(p - ez, cfg));
    return normalize_fd(n);
}

// ---------------------------------------------------------------------------
// Lighting (SM:5764-5861).
// ---------------------------------------------------------------------------
template <class T>
T preetham(V3<T> dr, V3<T> sun) {  // SM:5764-5774
    T Tt = T(2);
    T A = T(0.1787f) * Tt - T(1.4630f), B = T(-0.3554f) * Tt + T(0.4275f);
    T C = T(-0.0227f) * Tt + T(5.3251f), D = T(0.1206f) * Tt - T(2.5771f);
    T E = T(-0.0670f) * Tt + T(0.3703f);
    T theta = std::acos(m_clamp(dr.y, T(-1), T(1)));
    T gamma = std::acos(m_clamp(dot(dr, sun), T(-1), T(1)));
    return (T(1) + A * std::exp(B / m_max(T(0.1f), std::cos(theta)))) *
           (T(1) + C * std::exp(D * gamma) + E * std::pow(std::cos(gamma), T(2)));
}

template <class T>
V3<T> sky(V3<T> dr, V3<T> sun) {  // SM:5776-5787
    T sk = preetham(dr, sun);
    V3<T> col = vmix(V3<T>(T(0.004f), T(0.048f), T(0.253f)), V3<T>(T(0.8f), T(0.9f), T(1)), sk);
    T hd = (dr.y + T(1)) / T(2);
    T hs = (sun.y + T(1)) / T(2);
    col = vmix(V3<T>(T(1), T(0.85f), T(0.53f)) / T(4), col * T(1.2f), hd) * hs;
    col = (col - V3<T>(T(0.5), T(0.5), T(0.5))) * T(1.2f) + V3<T>(T(0.5), T(0.5), T(0.5));
    T disk = m_max((dot(dr, sun) - T(1)) / (T(1) - std::cos(T(0.05f))) + T(1), T(0));
    disk *= std::pow(hd, T(10));
    return col * T(0.8f) + V3<T>(disk, disk, disk) * T(100);
}

template <class T>
V3<T> background_gradient(V3<T> dir, const FptRenderConfig &cfg) {  // SM:5789-5794
    T t = m_clamp((dir.y + T(1)) / T(2), T(0), T(1));
    return vmix(load3<T>(&cfg.background_gradient[3]), load3<T>(&cfg.background_gradient[0]), t);
}

// IEEE half bits to float (exact).
inline float half_to_float(uint16_t h) {
    const int e = (h >> 10) & 31, m = h & 1023;
    const float mag = e == 0 ? std::ldexp(float(m), -24)
                             : (e == 31 ? (m ? std::numeric_limits<float>::quiet_NaN()
                                             : std::numeric_limits<float>::infinity())
                                        : std::ldexp(float(1024 + m), e - 25));
    return (h & 0x8000u) ? -mag : mag;
}

// Shading-language fract: x - floor(x), never 1.
template <class T>
T sl_fract(T x) {
    return m_min(x - std::floor(x), T(0x1.fffffep-1f));
}

template <class T>
V3<T> sample_hdri(V3<T> direction, const FptRenderConfig &cfg) {  // SM:5796-5821
    const uint32_t w = std::min(std::max(cfg.hdri_width, 1u), 32u);
    const uint32_t h = std::min(std::max(cfg.hdri_height, 1u), 16u);
    auto texel = [&](uint32_t x, uint32_t y) {
        const uint32_t i = (std::min(y, h - 1u) * w + std::min(x, w - 1u)) * 3u;
        return V3<T>(T(half_to_float(cfg.hdri_lut[i])), T(half_to_float(cfg.hdri_lut[i + 1])),
                     T(half_to_float(cfg.hdri_lut[i + 2])));
    };
    const V3<T> d = normalize(direction);
    const T u = sl_fract(std::atan2(d.z, d.x) / (T(2) * T(kPi)) + T(0.5));
    const T v = m_clamp(std::acos(m_clamp(d.y, T(-1), T(1))) / T(kPi), T(0), T(1));
    const T x = u * T(w) - T(0.5);
    const T y = v * T(h - 1u);
    const int x0i = int(std::floor(x));
    const uint32_t x0 = uint32_t((x0i % int(w) + int(w)) % int(w));
    const uint32_t x1 = (x0 + 1u) % w;
    const uint32_t y0 = uint32_t(std::min(std::max(int(std::floor(y)), 0), int(h - 1u)));
    const uint32_t y1 = std::min(y0 + 1u, h - 1u);
    const T fx = sl_fract(x), fy = sl_fract(y);
    return vmix(vmix(texel(x0, y0), texel(x1, y0), fx), vmix(texel(x0, y1), texel(x1, y1), fx), fy);
}

template <class T>
V3<T> environment(V3<T> view, const FptRenderConfig &cfg) {  // SM:5823-5844
    if (T(cfg.world[4]) < T(0.00001f)) return V3<T>();
    V3<T> ld = rotate_camera(V3<T>(T(0), T(0), T(1)), T(cfg.world[2]) * T(kPi) / T(180),
                             T(cfg.world[3]) * T(kPi) / T(180));
    V3<T> env;
    if (cfg.world[0] == 0.0f) {
        T ls = m_max(T(cfg.world[1]), T(0.0001f));
        T l = m_max((dot(view, ld) - T(1)) / (T(1) - std::cos(ls)) + T(1), T(0)) / ls * T(3);
        env = V3<T>(l, l, l);
    } else if (cfg.world[0] == 1.0f) {
        env = sky(view, ld);
    } else if (cfg.world[0] == 2.0f && cfg.hdri_enabled != 0u && cfg.hdri_width > 0u) {
        env = sample_hdri(rotate_camera(view, T(cfg.world[2]) * T(kPi) / T(180), T(0)), cfg);
    } else if (cfg.world[0] == 3.0f) {
        env = load3<T>(cfg.world_one_color);
    } else {
        env = V3<T>(T(1), T(1), T(1));
    }
    T power = T(cfg.world[4]);
    T contrast = cfg.world[0] == 3.0f ? T(1) : T(cfg.world[5]);
    V3<T> r = env * power - V3<T>(T(0.5), T(0.5), T(0.5)) * contrast + V3<T>(T(0.5), T(0.5), T(0.5));
    return V3<T>(m_max(r.x, T(0)), m_max(r.y, T(0)), m_max(r.z, T(0)));
}

template <class T>
V3<T> sun_contribution(V3<T> rp, UV xy, float seed, const Material<T> &mat, V3<T> n,
                       const FptRenderConfig &cfg) {  // SM:5846-5861
    T p1 = mat.roughness, p2 = T(1) - mat.translucency;
    V3<T> ld = rotate_camera(V3<T>(T(0), T(0), T(1)), T(cfg.sun[1]) * T(kPi) / T(180),
                             T(cfg.sun[2]) * T(kPi) / T(180));
    V3<T> rp0 = rp;
    rp = rp + n * T(0.001f);
    float k1 = mul_add_f32(seed, 5.0f, 1.0f);
    float k2 = mul_add_f32(seed, 3.0f, 5.0f);
    T h1 = T(hash13(xy.x, xy.y, k1));
    T h2 = T(hash13(xy.x, xy.y, k2));
    T dx = std::cos(h1 * T(2) * T(kPi)) * std::sqrt(h2) * T(cfg.sun[4]);
    T dy = std::sin(h1 * T(2) * T(kPi)) * std::sqrt(h2) * T(cfg.sun[4]);
    ld = rotate_camera(ld, dx, dy);
    int steps = 0;
    rp = march(ld, rp, int(cfg.render[1]), T(cfg.render[3]), T(0.0002f), cfg, &steps);
    if (tl_counters) tl_counters->shadow_steps += uint64_t(steps);
    if (length(rp0 - rp) > T(cfg.render[4]) * T(0.99f)) {
        T k = T(cfg.sun[3]) * m_max(dot(n, ld), T(0)) * p1 * p2;
        return load3<T>(cfg.sun_color) * k;
    }
    return V3<T>();
}

// One scatter event of renderPath (SM:5982-6013): the material's branch
// choice (r1 against translucency), then
//   reflective: dr = normalize(mix(reflect(dr, n), diffuse, roughness)),
//               Schlick Fresnel f0 + (1 - f0)(1 - cos)^5 times specular
//               against r2 picks the mirror direction (factor 1), otherwise
//               the colour is multiplied by the albedo;
//   refractive: eta = 1/ior entering (side 1), ior leaving (normal flipped);
//               total internal reflection (|refract|^2 < 1e-6 or not finite)
//               or r2 < Fresnel reflects, otherwise the ray refracts, the
//               side flips and the colour takes mix(1, rgb, 0.35)(1 - F/2).
// `diffuse` is randomVector's draw (used on the reflective branch only).
// Returns the new direction, the normal the bounce offset uses, the colour
// factor and the new side.
template <class T>
struct Scatter {
    V3<T> dr, n, factor;
    T side;
};

template <class T>
Scatter<T> scatter_event(V3<T> dr, V3<T> n, const Material<T> &mat, T side, T r1, T r2, V3<T> diffuse) {
    Scatter<T> out;
    out.factor = V3<T>(T(1), T(1), T(1));
    out.side = side;
    const T f0 = std::pow((mat.ior - T(1)) / (mat.ior + T(1)), T(2));
    if (r1 > mat.translucency) {
        const V3<T> metal = vreflect(dr, n);
        const T ct = m_clamp(dot(n, -dr), T(0), T(1));
        const T om = T(1) - ct;
        const T fres = f0 + (T(1) - f0) * (om * om * om * om * om);
        out.dr = normalize(vmix(metal, diffuse, mat.roughness));
        if (r2 < fres * mat.specular) out.dr = metal;
        else out.factor = mat.rgb;
        out.n = n;
        return out;
    }
    const T eta = side == T(1) ? T(1) / mat.ior : mat.ior;
    if (side != T(1)) n = -n;
    const T ct = m_clamp(dot(n, -dr), T(0), T(1));
    const T om = T(1) - ct;
    const T fres = f0 + (T(1) - f0) * (om * om * om * om * om);
    const V3<T> refr = vrefract(dr, n, eta);
    const V3<T> refl = vreflect(dr, n);
    if (dot(refr, refr) < T(0.000001f) || !std::isfinite(refr.x) || r2 < fres) {
        out.dr = normalize(refl);
    } else {
        out.dr = normalize(refr);
        out.side = -side;
        out.factor = vmix(V3<T>(T(1), T(1), T(1)), mat.rgb, T(0.35f)) * (T(1) - fres * T(0.5f));
    }
    out.n = n;
    return out;
}

// renderPath (SM:5926-6026).
template <class T>
V3<T> render_path(UV xy, uint32_t sample_idx, const FptRenderConfig &cfg) {
    float frame = float(sample_idx);
    V3<T> cam = load3<T>(cfg.camera_position);
    V3<T> rp = cam;
    volatile float aa_den = std::fmax(float(cfg.width), float(cfg.height));
    T aa = T(0.3f / aa_den);
    T jx, jy;
    random_point<T>(aa, xy, frame, jx, jy);
    V3<T> dr;
    if (!camera_ray<T>(T(xy.x) + jx, T(xy.y) + jy, cfg, dr)) return V3<T>();
    apply_dof<T>(xy, frame, cfg, rp, dr);
    int local_ni = int(cfg.render[1]);
    T side = T(1);
    V3<T> light, color(T(1), T(1), T(1));
    T sky_mask = T(0);
    V3<T> gcol = background_gradient(dr, cfg);
    int bounces = std::min(int(cfg.render[0]), 8);
    if (cfg.sdf_bounce_cap > 0u) bounces = std::min(bounces, int(cfg.sdf_bounce_cap));
    T maxd = T(cfg.render[4]);
    T maxd2 = maxd * maxd;
    T far = maxd * T(0.99f);
    T far2 = far * far;
    if (tl_counters) tl_counters->paths++;
    for (int i = 0; i < bounces; ++i) {
        int steps = 0;
        rp = march(dr, rp, local_ni, T(cfg.render[3]), T(0.0002f), cfg, &steps);
        if (tl_counters) (i == 0 ? tl_counters->primary_steps : tl_counters->secondary_steps) += uint64_t(steps);
        V3<T> tv = rp - cam;
        T travel2 = dot(tv, tv);
        if (i == 0 && travel2 > maxd2) sky_mask = T(1);
        local_ni = int(cfg.render[1] / (cfg.render[5] * 2.0f + 1.0f));
        if (travel2 > far2) {
            light = light + environment(dr, cfg);
            break;
        }
        Material<T> mat = user_material(rp, cfg);
        if (mat.emission > T(0.001f)) light = light + mat.rgb * mat.emission;
        V3<T> n = normal_at(rp, cfg);
        if (cfg.sun[0] == 1.0f) light = light + sun_contribution(rp, xy, frame, mat, n, cfg);
        T r1 = T(hash13(xy.x, xy.y, mul_add_f32(frame, 1.37f, float(i))));
        T r2 = T(hash13(xy.x, xy.y, mul_add_f32(frame, 7.91f, float(i))));
        // randomVector is drawn only on the reflective branch, as in SM:5986.
        const V3<T> diffuse =
            r1 > mat.translucency ? random_vector(n, xy, mul_add_f32(frame, 13.37f, float(i))) : V3<T>();
        const Scatter<T> sc = scatter_event(dr, n, mat, side, r1, r2, diffuse);
        dr = sc.dr;
        n = sc.n;
        side = sc.side;
        color = color * sc.factor;
        rp = rp + n * (T(0.001f) * m_sign(dot(dr, n)));
        if (cfg.sdf_russian_roulette != 0u && i + 1 < bounces && float(i + 1) >= cfg.sdf_rr_start) {
            T surv = m_clamp(m_max(color.x, m_max(color.y, color.z)),
                             m_clamp(T(cfg.sdf_rr_min_prob), T(0.01f), T(1)), T(1));
            float rk = mul_add_f32(frame, 19.19f, float(i) * 3.17f);
            T rr = T(hash13(xy.x, xy.y, rk));
            if (rr > surv) break;
            color = color / surv;
        }
    }
    V3<T> col = light * color;
    if (cfg.world[6] == 1.0f) col = vmix(col, gcol, sky_mask);
    return V3<T>(m_min(col.x, T(8)), m_min(col.y, T(8)), m_min(col.z, T(8)));
}

template <class T>
void render_rows(const FptRenderConfig &cfg, int threads, uint32_t y0, uint32_t y1, uint32_t s0, uint32_t s1,
                 double *accum) {
    const uint32_t W = cfg.width;
    const uint32_t tiles_x = (W + 15) / 16, tiles_y = (y1 - y0 + 15) / 16;
    const long ntiles = long(tiles_x) * long(tiles_y);
    Counters total;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
#endif
    {
        Counters local;
        tl_counters = &local;
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
        for (long t = 0; t < ntiles; ++t) {
            uint32_t tx = uint32_t(t % tiles_x), ty = uint32_t(t / tiles_x);
            for (uint32_t y = y0 + ty * 16; y < std::min(y1, y0 + ty * 16 + 16); ++y) {
                for (uint32_t x = tx * 16; x < std::min(W, tx * 16 + 16); ++x) {
                    UV uv = screen_uv(x, y, cfg);
                    V3<T> c;
                    for (uint32_t s = s0; s < s1; ++s) {
                        V3<T> sc = render_path<T>(uv, s, cfg);
                        T a = T(1) / T(s - s0 + 1);
                        c = vmix(c, sc, a);
                    }
                    double *o = accum + (size_t(y) * W + x) * 4;
                    o[0] = double(c.x); o[1] = double(c.y); o[2] = double(c.z); o[3] = 1.0;
                }
            }
        }
        tl_counters = nullptr;
#ifdef _OPENMP
#pragma omp critical
#endif
        total.add(local);
    }
    g_last = total;
}

template <class T>
void trace_primary(const FptRenderConfig &cfg, int threads, double *out) {
    const uint32_t W = cfg.width, H = cfg.height;
    const long n = long(W) * long(H);
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
#endif
    for (long i = 0; i < n; ++i) {
        uint32_t x = uint32_t(i % W), y = uint32_t(i / W);
        UV uv = screen_uv(x, y, cfg);
        V3<T> cam = load3<T>(cfg.camera_position);
        V3<T> dr;
        const bool seen = camera_ray<T>(T(uv.x), T(uv.y), cfg, dr);
        int steps = 0;
        V3<T> rp = seen ? march(dr, cam, int(cfg.render[1]), T(cfg.render[3]), T(0.0002f), cfg, &steps) : cam;
        V3<T> tv = rp - cam;
        T travel2 = dot(tv, tv);
        T far = T(cfg.render[4]) * T(0.99f);
        bool hit = seen && travel2 <= far * far;
        V3<T> nrm = hit ? normal_at(rp, cfg) : V3<T>();
        double *o = out + size_t(i) * 9;
        o[0] = double(std::sqrt(travel2));
        o[1] = double(rp.x); o[2] = double(rp.y); o[3] = double(rp.z);
        o[4] = double(nrm.x); o[5] = double(nrm.y); o[6] = double(nrm.z);
        o[7] = hit ? 1.0 : 0.0;
        o[8] = double(steps);
    }
}

#include "fpt_oracle_bounds.inc"
#include "fpt_oracle_voxel.inc"
#include "fpt_oracle_diag.inc"

}  // namespace

extern "C" {

int orc_render(const FptRenderConfig *cfg, int precision, int threads, uint32_t row_begin,
               uint32_t row_end, double *accum) {
    if (!cfg || !accum || cfg->width == 0 || cfg->height == 0) return 1;
    if (row_end > cfg->height || row_begin >= row_end) return 1;
    const uint32_t samples = std::min<uint32_t>(std::max<uint32_t>(cfg->samples, 1u), 512u);
    if (precision == 64) render_rows<double>(with_focus<double>(*cfg), threads, row_begin, row_end, 0u, samples, accum);
    else render_rows<float>(with_focus<float>(*cfg), threads, row_begin, row_end, 0u, samples, accum);
    return 0;
}

int orc_render_samples(const FptRenderConfig *cfg, int precision, int threads, uint32_t row_begin,
                       uint32_t row_end, uint32_t sample_begin, uint32_t sample_end, double *accum) {
    if (!cfg || !accum || cfg->width == 0 || cfg->height == 0) return 1;
    if (row_end > cfg->height || row_begin >= row_end || sample_begin >= sample_end) return 1;
    if (precision == 64)
        render_rows<double>(with_focus<double>(*cfg), threads, row_begin, row_end, sample_begin, sample_end, accum);
    else
        render_rows<float>(with_focus<float>(*cfg), threads, row_begin, row_end, sample_begin, sample_end, accum);
    return 0;
}

// Leaf refinement probe: mode 1 secant/bisection, 2 restricted trace, 3 fixed
// DE; ray[0..5] origin and direction, ray[6..7] the interval, ray[8..13] the
// cell centre and half extent.  Returns 1 when refined (out = distance).
int orc_leaf_refine(const FptRenderConfig *cfg, int precision, uint32_t mode, const double *ray, double *out) {
    if (!cfg || !ray || !out) return 0;
    auto run = [&](auto zero) {
        using T = decltype(zero);
        const V3<T> o{T(ray[0]), T(ray[1]), T(ray[2])}, d{T(ray[3]), T(ray[4]), T(ray[5])};
        const V3<T> c{T(ray[8]), T(ray[9]), T(ray[10])}, h{T(ray[11]), T(ray[12]), T(ray[13])};
        T t = T(0);
        bool ok = false;
        if (mode == 1u) ok = refine_secant_bisection<T>(o, d, T(ray[6]), T(ray[7]), *cfg, t);
        else if (mode == 2u) ok = refine_restricted_trace<T>(o, d, T(ray[6]), T(ray[7]), c, h, *cfg, t);
        else if (mode == 3u) ok = refine_fixed_de<T>(o, d, T(ray[6]), T(ray[7]), *cfg, t);
        *out = double(t);
        return ok ? 1 : 0;
    };
    return precision == 64 ? run(0.0) : run(0.0f);
}

int orc_environment(const FptRenderConfig *cfg, int precision, const double *dirs, size_t n, double *out) {
    if (!cfg || (n && (!dirs || !out))) return 1;
    for (size_t i = 0; i < n; ++i) {
        const double *d = dirs + 3 * i;
        if (precision == 64) {
            const V3<double> e = environment(V3<double>(d[0], d[1], d[2]), *cfg);
            out[3 * i] = e.x; out[3 * i + 1] = e.y; out[3 * i + 2] = e.z;
        } else {
            const V3<float> e = environment(V3<float>(float(d[0]), float(d[1]), float(d[2])), *cfg);
            out[3 * i] = e.x; out[3 * i + 1] = e.y; out[3 * i + 2] = e.z;
        }
    }
    return 0;
}

double orc_focus_estimate(const FptRenderConfig *cfg, int precision) {
    if (!cfg) return -1.0;
    return precision == 64 ? focus_estimate<double>(*cfg) : double(focus_estimate<float>(*cfg));
}

int orc_trace_primary(const FptRenderConfig *cfg, int precision, int threads, double *out) {
    if (!cfg || !out || cfg->width == 0 || cfg->height == 0) return 1;
    if (precision == 64) trace_primary<double>(*cfg, threads, out);
    else trace_primary<float>(*cfg, threads, out);
    return 0;
}

int orc_map_distance(const FptRenderConfig *cfg, int precision, const double *pts, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i) {
        if (precision == 64)
            out[i] = map_sdf(V3<double>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), *cfg);
        else
            out[i] = double(map_sdf(V3<float>(float(pts[3 * i]), float(pts[3 * i + 1]), float(pts[3 * i + 2])), *cfg));
    }
    return 0;
}

int orc_normal(const FptRenderConfig *cfg, int precision, const double *pts, size_t n, double *normals) {
    for (size_t i = 0; i < n; ++i) {
        if (precision == 64) {
            V3<double> r = normal_at(V3<double>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), *cfg);
            normals[3 * i] = r.x; normals[3 * i + 1] = r.y; normals[3 * i + 2] = r.z;
        } else {
            V3<float> r = normal_at(V3<float>(float(pts[3 * i]), float(pts[3 * i + 1]), float(pts[3 * i + 2])), *cfg);
            normals[3 * i] = r.x; normals[3 * i + 1] = r.y; normals[3 * i + 2] = r.z;
        }
    }
    return 0;
}

// Pixel (x, y) -> screen uv -> camera ray, un-jittered: out = {u, v, dx,
// dy, dz, seen} per pixel.
int orc_camera_ray(const FptRenderConfig *cfg, int precision, const uint32_t *xy, size_t n, double *out) {
    if (!cfg || cfg->width == 0 || cfg->height == 0) return 1;
    for (size_t i = 0; i < n; ++i) {
        const UV uv = screen_uv(xy[2 * i], xy[2 * i + 1], *cfg);
        double *o = out + 6 * i;
        o[0] = uv.x;
        o[1] = uv.y;
        if (precision == 64) {
            V3<double> d;
            o[5] = camera_ray<double>(uv.x, uv.y, *cfg, d) ? 1.0 : 0.0;
            o[2] = d.x; o[3] = d.y; o[4] = d.z;
        } else {
            V3<float> d;
            o[5] = camera_ray<float>(uv.x, uv.y, *cfg, d) ? 1.0 : 0.0;
            o[2] = d.x; o[3] = d.y; o[4] = d.z;
        }
    }
    return 0;
}

// The scene's march from rays {ox, oy, oz, dx, dy, dz}: out = {x, y, z,
// steps} per ray (threshold render[3], LOD falloff 0.0002, render[1] steps).
int orc_march(const FptRenderConfig *cfg, int precision, const double *rays, size_t n, double *out) {
    if (!cfg) return 1;
    for (size_t i = 0; i < n; ++i) {
        const double *r = rays + 6 * i;
        int steps = 0;
        double *o = out + 4 * i;
        if (precision == 64) {
            V3<double> q = march(V3<double>(r[3], r[4], r[5]), V3<double>(r[0], r[1], r[2]), int(cfg->render[1]),
                                 double(cfg->render[3]), double(0.0002f), *cfg, &steps);
            o[0] = q.x; o[1] = q.y; o[2] = q.z;
        } else {
            V3<float> q = march(V3<float>(float(r[3]), float(r[4]), float(r[5])),
                                V3<float>(float(r[0]), float(r[1]), float(r[2])), int(cfg->render[1]),
                                cfg->render[3], 0.0002f, *cfg, &steps);
            o[0] = q.x; o[1] = q.y; o[2] = q.z;
        }
        o[3] = steps;
    }
    return 0;
}

int orc_program_surface(const FptRenderConfig *cfg, int precision, const double *pts, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i) {
        if (precision == 64) {
            V3<double> g;
            double d = program_surface<double>(geometry_program(*cfg), V3<double>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), g);
            out[4 * i] = d; out[4 * i + 1] = g.x; out[4 * i + 2] = g.y; out[4 * i + 3] = g.z;
        } else {
            V3<float> g;
            float d = program_surface<float>(geometry_program(*cfg),
                                             V3<float>(float(pts[3 * i]), float(pts[3 * i + 1]), float(pts[3 * i + 2])), g);
            out[4 * i] = d; out[4 * i + 1] = g.x; out[4 * i + 2] = g.y; out[4 * i + 3] = g.z;
        }
    }
    return 0;
}

int orc_leaf_orbit(uint32_t opcode, const float data[4], int precision, const double point[3], double out[9]) {
    if (precision == 64) {
        LeafResult<double> r;
        if (!eval_leaf<double, double>(opcode, V3<double>(point[0], point[1], point[2]), data, r)) return 1;
        out[0] = r.d;
        out[1] = r.radius; out[2] = r.derivative; out[3] = r.iterations; out[4] = r.escaped;
        out[5] = r.z.x; out[6] = r.z.y; out[7] = r.z.z; out[8] = r.trap;
    } else {
        LeafResult<float> r;
        if (!eval_leaf<float, float>(opcode, V3<float>(float(point[0]), float(point[1]), float(point[2])), data, r)) return 1;
        out[0] = r.d;
        out[1] = r.radius; out[2] = r.derivative; out[3] = r.iterations; out[4] = r.escaped;
        out[5] = r.z.x; out[6] = r.z.y; out[7] = r.z.z; out[8] = r.trap;
    }
    return 0;
}

void orc_mandelbulb_step(double s[5], double power) {
    V3<double> z(s[0], s[1], s[2]);
    double dr = s[4];
    mandelbulb_kernel<double, double>(z, s[3], dr, power);
    s[0] = z.x; s[1] = z.y; s[2] = z.z; s[3] = length(z); s[4] = dr;
}

void orc_mandelbox_step(double s[5], double scale, double min_r2, double fixed_r2) {
    V3<double> z(s[0], s[1], s[2]);
    double dr = s[4];
    mandelbox_kernel<double, double>(z, dr, scale, min_r2, fixed_r2);
    s[0] = z.x; s[1] = z.y; s[2] = z.z; s[3] = length(z); s[4] = dr;
}

int orc_program_bounds(const FptRenderConfig *cfg, int precision, const double center[3], const double half[3],
                       double out[4]) {
    if (!cfg || !center || !half || !out) return 1;
    if (precision == 64) {
        BoundsT<double> b = program_bounds<double>(V3<double>(center[0], center[1], center[2]),
                                                   V3<double>(half[0], half[1], half[2]), *cfg);
        out[0] = b.minimum; out[1] = b.maximum; out[2] = b.lipschitz; out[3] = b.certified ? 1.0 : 0.0;
    } else {
        BoundsT<float> b = program_bounds<float>(V3<float>(float(center[0]), float(center[1]), float(center[2])),
                                                 V3<float>(float(half[0]), float(half[1]), float(half[2])), *cfg);
        out[0] = b.minimum; out[1] = b.maximum; out[2] = b.lipschitz; out[3] = b.certified ? 1.0 : 0.0;
    }
    return 0;
}

int orc_classify_cell(const FptRenderConfig *cfg, int precision, const double center[3], const double half[3],
                      double out[7]) {
    if (!cfg || !center || !half || !out) return 1;
    CellClassT r;
    if (precision == 64) {
        V3<double> c{center[0], center[1], center[2]};
        r = classify_cell<double>(c, V3<double>(half[0], half[1], half[2]), distance_sdf<double>(c, *cfg), *cfg);
    } else {
        V3<float> c{float(center[0]), float(center[1]), float(center[2])};
        r = classify_cell<float>(c, V3<float>(float(half[0]), float(half[1]), float(half[2])),
                                 distance_sdf<float>(c, *cfg), *cfg);
    }
    out[0] = r.interval_min; out[1] = r.interval_max; out[2] = r.certified; out[3] = r.legacy;
    out[4] = r.lipschitz; out[5] = r.interval; out[6] = double(r.cell_class);
    return 0;
}

int orc_kifs_sample(const FptRenderConfig *cfg, int precision, const double *points, size_t n, double *out) {
    if (!cfg || (n && (!points || !out))) return 1;
    for (size_t i = 0; i < n; ++i) {
        const double *q = points + 3 * i;
        if (precision == 64) {
            kifs_field_sample<double>(V3<double>(q[0], q[1], q[2]), *cfg, 1, out + 4 * i);
        } else {
            float o[4];
            kifs_field_sample<float>(V3<float>{float(q[0]), float(q[1]), float(q[2])}, *cfg, 1, o);
            for (int k = 0; k < 4; ++k) out[4 * i + k] = o[k];
        }
    }
    return 0;
}

int orc_present(const FptRenderConfig *cfg, int precision, const float *accum, uint8_t *rgba) {
    if (!cfg || !accum || !rgba || cfg->width == 0 || cfg->height == 0) return 1;
    if (precision == 64) present<double>(*cfg, accum, rgba);
    else present<float>(*cfg, accum, rgba);
    return 0;
}

int orc_diagnostic(const FptRenderConfig *cfg, uint32_t mode, uint32_t target_bounce, float max_distance,
                   int precision, int threads, uint8_t *rgba) {
    if (!cfg || !rgba || cfg->width == 0 || cfg->height == 0 || mode > FPT_DIAGNOSTIC_SDF_BOUNCE_CONTRIBUTION)
        return 1;
    const bool voxel = cfg->renderer_backend == FPT_RENDERER_VOXEL;
    VoxTables t;
    float focus = 0.0f;
    if (voxel) {
        if (!vx_build(*cfg, threads, t)) return 1;
        focus = vx_focus(*cfg, t);
    }
    const FptRenderConfig sdf_cfg = voxel ? *cfg : (precision == 64 ? with_focus<double>(*cfg) : with_focus<float>(*cfg));
    const uint32_t W = cfg->width, H = cfg->height;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 2) num_threads(threads > 0 ? threads : omp_get_max_threads())
#endif
    for (long y = 0; y < long(H); ++y) {
        for (uint32_t x = 0; x < W; ++x) {
            UV uv = screen_uv(x, uint32_t(y), *cfg);
            uint8_t *o = rgba + (size_t(H - 1 - uint32_t(y)) * W + x) * 4;
            if (voxel) {
                V3<float> c = voxel_diagnostic(uv, *cfg, t, focus, mode, max_distance);
                store_rgba8(o, c.x, c.y, c.z);
            } else if (precision == 64) {
                V3<double> c = sdf_diagnostic<double>(uv, sdf_cfg, mode, target_bounce, max_distance);
                store_rgba8(o, c.x, c.y, c.z);
            } else {
                V3<float> c = sdf_diagnostic<float>(uv, sdf_cfg, mode, target_bounce, max_distance);
                store_rgba8(o, c.x, c.y, c.z);
            }
        }
    }
    return 0;
}

// randomVector (SM:557-568) about normals[3i..3i+2] with keys
// keys[3i..3i+2] = {uv.x, uv.y, seed} (float32): out[3i..3i+2].
int orc_random_vector(int precision, const double *normals, const float *keys, size_t n, double *out) {
    if (!normals || !keys || !out) return 1;
    for (size_t i = 0; i < n; ++i) {
        const double *nv = normals + 3 * i;
        const UV xy{keys[3 * i], keys[3 * i + 1]};
        if (precision == 64) {
            const V3<double> d = random_vector(V3<double>(nv[0], nv[1], nv[2]), xy, keys[3 * i + 2]);
            out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
        } else {
            const V3<float> d =
                random_vector(V3<float>(float(nv[0]), float(nv[1]), float(nv[2])), xy, keys[3 * i + 2]);
            out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
        }
    }
    return 0;
}

// One scatter event (scatter_event): in = {dr[3], n[3], rgb[3], roughness,
// specular, translucency, ior, side, r1, r2, diffuse[3]} (19 doubles),
// out = {dr[3], n[3], factor[3], side} (10 doubles).
int orc_scatter(int precision, const double *in, double *out) {
    if (!in || !out) return 1;
    auto run = [&](auto zero) {
        using T = decltype(zero);
        Material<T> m;
        m.rgb = V3<T>(T(in[6]), T(in[7]), T(in[8]));
        m.roughness = T(in[9]);
        m.specular = T(in[10]);
        m.translucency = T(in[11]);
        m.ior = T(in[12]);
        const Scatter<T> sc = scatter_event(V3<T>(T(in[0]), T(in[1]), T(in[2])), V3<T>(T(in[3]), T(in[4]), T(in[5])),
                                            m, T(in[13]), T(in[14]), T(in[15]),
                                            V3<T>(T(in[16]), T(in[17]), T(in[18])));
        const double v[10] = {double(sc.dr.x), double(sc.dr.y), double(sc.dr.z), double(sc.n.x), double(sc.n.y),
                              double(sc.n.z), double(sc.factor.x), double(sc.factor.y), double(sc.factor.z),
                              double(sc.side)};
        std::memcpy(out, v, sizeof(v));
    };
    if (precision == 64) run(0.0);
    else run(0.0f);
    return 0;
}

uint32_t orc_hash13_bits(float x, float y, float z) { return hash13bits(x, y, z); }
float orc_hash13(float x, float y, float z) { return hash13(x, y, z); }

void orc_last_counters(uint64_t out[9]) {
    out[7] = g_last.far_steps;
    out[8] = g_last.far_orbit_iterations;
    out[0] = g_last.distance_evals;
    out[1] = g_last.orbit_iterations;
    out[2] = g_last.primary_steps;
    out[3] = g_last.secondary_steps;
    out[4] = g_last.shadow_steps;
    out[5] = g_last.normal_evals;
    out[6] = g_last.paths;
}

}  // extern "C"
