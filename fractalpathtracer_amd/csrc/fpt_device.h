// Device-side scene evaluation for gfx950: vector helpers, the reference's
// RNG and camera, the typed SDF program interpreter with fractal leaves, the
// analytic program gradient, march, normals, lighting and the path tracer.
//
// Reference semantics: shaders/Shaders.metal (SM) and
// src/mandelbulber/orbit.rs (ORB); line numbers are cited per function.
//
// Design notes (MI355X):
//  * All geometry is FP32.  RNG keys and the pixel uv use explicitly rounded
//    intrinsics (__fmul_rn/__fadd_rn/__fdiv_rn) so they are bit-identical to
//    an IEEE CPU computation: the hash keys on float bit patterns.
//  * The integer-power Mandelbulb step avoids the seven transcendentals of
//    the triplex formula (asin, atan2, pow, 2 sin, 2 cos): (cos t + i sin t)^p
//    and (cos f + i sin f)^p are raised by repeated complex squaring from
//    cos/sin that are ratios of the orbit coordinates.  It is the same
//    function in exact arithmetic; only VALU adds/muls, one sqrt and one
//    reciprocal remain per iteration.
//  * Leaves are templated on the scalar type so the analytic normal is the
//    forward-mode derivative (Dual) of exactly the function the march uses.
#pragma once

#ifndef __HIPCC_RTC__  // the runtime compiler provides these itself
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "../../include/fpt_abi.h"

namespace fpt {

constexpr float kPi = 3.14159265359f;  // SM:4
constexpr float kInf = 1.0e20f;        // SM:5

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float k) { return {a.x * k, a.y * k, a.z * k}; }
__device__ __forceinline__ f3 operator/(f3 a, float k) { return {a.x / k, a.y / k, a.z / k}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 vabs(f3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ f3 vmix(f3 a, f3 b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ f3 vmax0(f3 a) { return {fmaxf(a.x, 0.f), fmaxf(a.y, 0.f), fmaxf(a.z, 0.f)}; }
__device__ __forceinline__ f3 vmin(f3 a, float k) { return {fminf(a.x, k), fminf(a.y, k), fminf(a.z, k)}; }
__device__ __forceinline__ f3 reflect(f3 i, f3 n) { return i - n * (2.0f * dot(n, i)); }
__device__ __forceinline__ f3 refract(f3 i, f3 n, float eta) {
    float ni = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - ni * ni);
    if (k < 0.0f) return {0.f, 0.f, 0.f};
    return i * eta - n * (eta * ni + sqrtf(k));
}
__device__ __forceinline__ float signf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
__device__ __forceinline__ f3 ld3(const float *v) { return {v[0], v[1], v[2]}; }
__device__ __forceinline__ bool finite3(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// ---------------------------------------------------------------------------
// Forward-mode dual number: value + gradient wrt the 3 point coordinates.
// ---------------------------------------------------------------------------
struct Df {
    float v, a, b, c;
};
__device__ __forceinline__ Df dconst(float x) { return {x, 0.f, 0.f, 0.f}; }
__device__ __forceinline__ Df operator+(Df x, Df y) { return {x.v + y.v, x.a + y.a, x.b + y.b, x.c + y.c}; }
__device__ __forceinline__ Df operator-(Df x, Df y) { return {x.v - y.v, x.a - y.a, x.b - y.b, x.c - y.c}; }
__device__ __forceinline__ Df operator-(Df x) { return {-x.v, -x.a, -x.b, -x.c}; }
__device__ __forceinline__ Df operator*(Df x, Df y) {
    return {x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b, x.c * y.v + x.v * y.c};
}
__device__ __forceinline__ Df operator*(Df x, float k) { return {x.v * k, x.a * k, x.b * k, x.c * k}; }
__device__ __forceinline__ Df operator/(Df x, Df y) {
    float inv = 1.0f / y.v;
    float q = x.v * inv;
    return {q, (x.a - q * y.a) * inv, (x.b - q * y.b) * inv, (x.c - q * y.c) * inv};
}

// Scalar traits: the same leaf code runs on float and Df.
__device__ __forceinline__ float val(float x) { return x; }
__device__ __forceinline__ float val(Df x) { return x.v; }
__device__ __forceinline__ float s_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ Df s_sqrt(Df x) {
    float r = sqrtf(x.v);
    float k = 0.5f / r;
    return {r, x.a * k, x.b * k, x.c * k};
}
__device__ __forceinline__ float s_rcp(float x) { return 1.0f / x; }
__device__ __forceinline__ Df s_rcp(Df x) {
    float r = 1.0f / x.v;
    float k = -r * r;
    return {r, x.a * k, x.b * k, x.c * k};
}
__device__ __forceinline__ float s_log(float x) { return logf(x); }
__device__ __forceinline__ Df s_log(Df x) {
    float k = 1.0f / x.v;
    return {logf(x.v), x.a * k, x.b * k, x.c * k};
}
__device__ __forceinline__ float s_sin(float x) { return sinf(x); }
__device__ __forceinline__ Df s_sin(Df x) {
    float s, c;
    sincosf(x.v, &s, &c);
    return {s, x.a * c, x.b * c, x.c * c};
}
__device__ __forceinline__ float s_cos(float x) { return cosf(x); }
__device__ __forceinline__ Df s_cos(Df x) {
    float s, c;
    sincosf(x.v, &s, &c);
    return {c, -x.a * s, -x.b * s, -x.c * s};
}
__device__ __forceinline__ float s_asin(float x) { return asinf(x); }
__device__ __forceinline__ Df s_asin(Df x) {
    float k = 1.0f / sqrtf(1.0f - x.v * x.v);
    return {asinf(x.v), x.a * k, x.b * k, x.c * k};
}
__device__ __forceinline__ float s_atan2(float y, float x) { return atan2f(y, x); }
__device__ __forceinline__ Df s_atan2(Df y, Df x) {
    float den = 1.0f / (x.v * x.v + y.v * y.v);
    float ky = x.v * den, kx = -y.v * den;
    return {atan2f(y.v, x.v), ky * y.a + kx * x.a, ky * y.b + kx * x.b, ky * y.c + kx * x.c};
}
__device__ __forceinline__ float s_pow(float x, float e) { return powf(x, e); }
__device__ __forceinline__ Df s_pow(Df x, float e) {
    float p = powf(x.v, e);
    float k = e * powf(x.v, e - 1.0f);
    return {p, x.a * k, x.b * k, x.c * k};
}
__device__ __forceinline__ float s_abs(float x) { return fabsf(x); }
__device__ __forceinline__ Df s_abs(Df x) { return x.v < 0.f ? -x : x; }
__device__ __forceinline__ float s_max(float x, float y) { return fmaxf(x, y); }
__device__ __forceinline__ Df s_max(Df x, Df y) { return x.v < y.v ? y : x; }
__device__ __forceinline__ float s_min(float x, float y) { return fminf(x, y); }
__device__ __forceinline__ Df s_min(Df x, Df y) { return y.v < x.v ? y : x; }
__device__ __forceinline__ float s_scale(float x, float k) { return x * k; }
__device__ __forceinline__ Df s_scale(Df x, float k) { return x * k; }

// Native single-instruction reciprocal / square root (v_rcp_f32, v_sqrt_f32,
// about 1 ulp) for the orbit loop, where the correctly rounded sequences would
// cost ~10 instructions each.
__device__ __forceinline__ float f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ Df f_rcp(Df x) {
    float r = __builtin_amdgcn_rcpf(x.v);
    float k = -r * r;
    return {r, x.a * k, x.b * k, x.c * k};
}
__device__ __forceinline__ float f_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ Df f_sqrt(Df x) {
    float r = __builtin_amdgcn_sqrtf(x.v);
    float k = 0.5f * __builtin_amdgcn_rcpf(r);
    return {r, x.a * k, x.b * k, x.c * k};
}

template <class S>
struct v3 {
    S x, y, z;
};
template <class S> __device__ __forceinline__ v3<S> add(v3<S> a, v3<S> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class S> __device__ __forceinline__ S dot3(v3<S> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

// ---------------------------------------------------------------------------
// RNG (SM:541-576): MurmurHash2-style mix over float32 bit patterns.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hash13bits(float x, float y, float z) {
    uint32_t h = 1190494759u;
    h = (h ^ __float_as_uint(x)) * 0x5bd1e995u;
    h = (h ^ __float_as_uint(y)) * 0x5bd1e995u;
    h = (h ^ __float_as_uint(z)) * 0x5bd1e995u;
    h ^= h >> 13u;
    h *= 0x5bd1e995u;
    h ^= h >> 15u;
    return h;
}
__device__ __forceinline__ float hash13(float x, float y, float z) {
    // (h & 0xffffff) / 2^24 is exact in float32.
    return float(hash13bits(x, y, z) & 0x00ffffffu) * (1.0f / 16777216.0f);
}
// float32 a*b + c rounded after the product (the reference source order).
__device__ __forceinline__ float madd_rn(float a, float b, float c) { return __fadd_rn(__fmul_rn(a, b), c); }

struct UV {
    float x, y;
};

// randomVector (SM:557-568).
__device__ __forceinline__ f3 random_vector(f3 normal, UV xy, float seed) {
    float k = __fmul_rn(seed, 2.0f);
    float h1 = hash13(xy.x, xy.y, __fadd_rn(k, 0.0f));
    float h2 = hash13(xy.x, xy.y, __fadd_rn(k, 1.0f));
    f3 n = normalize(normal);
    f3 uu = normalize(cross(n, mk(0.f, 1.f, 1.f)));
    f3 vv = cross(uu, n);
    float ra = sqrtf(h2);
    float s, c;
    sincosf(kPi * 2.0f * h1, &s, &c);
    float rz = sqrtf(fmaxf(1.0f - h2, 0.0f));
    return normalize(uu * (ra * c) + vv * (ra * s) + n * rz);
}

// randomPoint (SM:570-576).
__device__ __forceinline__ void random_point(float power, UV xy, float seed, float &ox, float &oy) {
    float k = __fmul_rn(seed, 2.0f);
    float h1 = hash13(xy.x, xy.y, __fadd_rn(k, 0.0f));
    float h2 = hash13(xy.x, xy.y, __fadd_rn(k, 1.0f));
    float r = 2.0f * h1 * kPi;
    float r2 = 2.0f * h2 * kPi;
    float s, c;
    sincosf(r, &s, &c);
    float m = sqrtf(r2) * power;
    ox = c * m;
    oy = s * m;
}

// ---------------------------------------------------------------------------
// Camera (SM:460-471, 6847-6855).
// ---------------------------------------------------------------------------
__device__ __forceinline__ f3 rotate_camera(f3 v, float yaw, float pitch) {
    float sp, cp, sy, cy;
    sincosf(pitch, &sp, &cp);
    sincosf(yaw, &sy, &cy);
    v = mk(v.x, v.z * sp + v.y * cp, v.z * cp - v.y * sp);
    return mk(v.x * cy + v.z * sy, v.y, -v.x * sy + v.z * cy);
}
__device__ __forceinline__ f3 rotate_camera_roll(f3 v, float yaw, float pitch, float roll) {
    float s, c;
    sincosf(roll, &s, &c);
    return rotate_camera(mk(c * v.x - s * v.y, s * v.x + c * v.y, v.z), yaw, pitch);
}

__device__ __forceinline__ UV screen_uv(uint32_t x, uint32_t y, uint32_t w, uint32_t h) {
    float fw = float(w), fh = float(h);
    float u = __fsub_rn(__fdiv_rn(__fadd_rn(float(x), 0.5f), fw), 0.5f);
    float v = __fsub_rn(__fdiv_rn(__fadd_rn(float(y), 0.5f), fh), 0.5f);
    return {__fmul_rn(u, __fdiv_rn(fw, fh)), v};
}

// ---------------------------------------------------------------------------
// Fractal leaves.
// ---------------------------------------------------------------------------
template <class S>
struct Leaf {
    S d;
    float radius, derivative, trap;
    int iterations;
};

// |z| robust to float32 overflow of |z|^2 (escaping Mandelbulb orbits pass
// |z| ~ 1e20 one iteration after |p| ~ 235): the scaled form for lanes whose
// |z|^2 is not finite.
template <class S>
__device__ __forceinline__ S orbit_radius_slow(v3<S> z, S r2) {
    float m = fmaxf(fabsf(val(z.x)), fmaxf(fabsf(val(z.y)), fabsf(val(z.z))));
    if (!isfinite(m)) return r2;
    float im = 1.0f / m;
    v3<S> q = {s_scale(z.x, im), s_scale(z.y, im), s_scale(z.z, im)};
    return s_scale(s_sqrt(dot3(q)), m);
}
template <class S>
__device__ __forceinline__ S orbit_radius(v3<S> z) {
    S r2 = dot3(z);
    S r = f_sqrt(r2);
    const bool slow = !isfinite(val(r2));
    // A wave-uniform test: the common case (no lane overflows) costs one
    // ballot instead of a divergent branch's exec-mask setup per call.
    if (__builtin_expect(__any(slow), 0)) {
        if (slow) r = orbit_radius_slow(z, r2);
    }
    return r;
}

// (c + i s)^n for integer n >= 1 by binary exponentiation.
template <class S>
__device__ __forceinline__ void cpow_int(S c, S s, int n, S &oc, S &os) {
    S rc = c, rs = s;
    bool have = false;
    S ac = c, as = s;
    while (true) {
        if (n & 1) {
            if (!have) {
                rc = ac;
                rs = as;
                have = true;
            } else {
                S t = rc * ac - rs * as;
                rs = rc * as + rs * ac;
                rc = t;
            }
        }
        n >>= 1;
        if (!n) break;
        S t = ac * ac - as * as;
        as = (ac * as) * 2.0f;
        ac = t;
    }
    oc = rc;
    os = rs;
}
__device__ __forceinline__ Df operator*(Df x, int k) { return x * float(k); }

template <class S>
__device__ __forceinline__ S ipow_int(S x, int n) {  // x^n, n >= 0
    S r = x;
    bool have = false;
    S b = x;
    while (n) {
        if (n & 1) {
            r = have ? r * b : b;
            have = true;
        }
        n >>= 1;
        if (n) b = b * b;
    }
    if (!have) {
        if constexpr (sizeof(S) == sizeof(float)) return S(1.0f);
        else return dconst(1.0f);
    }
    return r;
}

template <class S> __device__ __forceinline__ S sconst(float x);
template <> __device__ __forceinline__ float sconst<float>(float x) { return x; }
template <> __device__ __forceinline__ Df sconst<Df>(float x) { return dconst(x); }

// One triplex power step z <- r^p (cos t cos f, cos t sin f, sin t),
// dr <- p r^(p-1) dr + 1 (ORB:766-778).  P > 0: the integer power is a
// compile-time constant (straight-line complex squarings); P == 0 with
// ip > 0: runtime integer power; ip == 0: trigonometric form.
template <class S, int P>
__device__ __forceinline__ void bulb_step(v3<S> &z, S r, S &dr, float power, int ip) {
    if (P > 0 || ip > 0) {
        const int n = P > 0 ? P : ip;
        S rho2 = z.x * z.x + z.y * z.y;
        S rho = f_sqrt(rho2);
        S ir = f_rcp(r);
        S st = z.z * ir;
        S ct = rho * ir;
        S cp, sp;
        if (val(rho) > 0.0f) {
            S irho = f_rcp(rho);
            cp = z.x * irho;
            sp = z.y * irho;
        } else {  // atan2(0, +-0) = 0 or pi: cos = sign(x) (x = +0 -> 1)
            cp = sconst<S>(signbit(val(z.x)) ? -1.0f : 1.0f);
            sp = sconst<S>(0.0f);
        }
        S cpt, spt, cpp, spp;
        cpow_int(ct, st, n, cpt, spt);
        cpow_int(cp, sp, n, cpp, spp);
        S rp = ipow_int(r, n - 1);
        dr = rp * dr * power + sconst<S>(1.0f);
        rp = rp * r;
        S k = cpt * rp;
        z = {k * cpp, k * spp, spt * rp};
    } else {
        S u = z.z / r;
        S theta = s_asin(u) * power;
        S phi = s_atan2(z.y, z.x) * power;
        S rp = s_pow(r, power - 1.0f);
        dr = rp * dr * power + sconst<S>(1.0f);
        rp = rp * r;
        S ctt = s_cos(theta);
        z = {ctt * s_cos(phi) * rp, ctt * s_sin(phi) * rp, s_sin(theta) * rp};
    }
}

// Logarithmic finalizer of the Mandelbulb orbit (ORB:654-660).
template <class S>
__device__ __forceinline__ S bulb_finish(S r, S dr) {
    if (!(val(dr) > 0.0f)) return r;
    if (val(r) > 1.0f) return (r * s_log(r) / dr) * 0.5f;
    return sconst<S>(0.0f);
}

template <class S, int P>
__device__ __forceinline__ Leaf<S> bulb_orbit(v3<S> p, float power, int ip, int iters, float bailout,
                                              int *orbit_counter) {
    v3<S> z = p;
    S r = orbit_radius(z);
    S dr = sconst<S>(1.0f);
    float trap = val(r);
    int done = 0;
    for (int i = 0; i < iters; ++i) {
        v3<S> old = z;
        bulb_step<S, P>(z, r, dr, power, ip);
        z = add(z, p);
        r = orbit_radius(z);
        done = i + 1;
        if (!isfinite(val(r))) {
            z = old;
            r = orbit_radius(z);
            break;
        }
        trap = fminf(trap, val(r));
        if (val(r) > bailout) break;
    }
    if (orbit_counter) *orbit_counter += done;
    Leaf<S> out;
    out.d = bulb_finish(r, dr);
    out.radius = val(r);
    out.derivative = val(dr);
    out.trap = trap;
    out.iterations = done;
    return out;
}

// Mandelbulb leaf: orbit loop (ORB:530-585) with the triplex kernel
// (ORB:766-778), +c (mandelbulber.rs:1421-1440) and the Logarithmic
// finalizer (ORB:654-660).  Power 8 (the common case) gets a specialised
// loop; the dispatch is on a config value, so it is wave-uniform.
template <class S>
__device__ Leaf<S> leaf_mandelbulb(v3<S> p, const float *data, int *orbit_counter) {
    const float power = data[0];
    const int iters = int(data[1]);
    const float bailout = data[2];
    const int ipw = int(power);
    const int ip = (float(ipw) == power && ipw >= 1 && ipw <= 32) ? ipw : 0;
    if (ip == 8) return bulb_orbit<S, 8>(p, power, 8, iters, bailout, orbit_counter);
    return bulb_orbit<S, 0>(p, power, ip, iters, bailout, orbit_counter);
}

// Mandelbox leaf: ORB:1636-1687 (fold limit 1, value 2, no rotations,
// no offset), orbit loop ORB:530-585, bailout 100, Linear finalizer.
template <class S>
__device__ Leaf<S> leaf_mandelbox(v3<S> p, const float *data, int *orbit_counter) {
    const float scale = data[0];
    const int iters = int(data[1]);
    const float min_r2 = data[2] * data[2];
    const float fixed_r2 = data[3] * data[3];
    const float min_factor = fixed_r2 / min_r2;
    const float abs_scale = fabsf(scale);
    v3<S> z = p;
    S r = orbit_radius(z);
    S dr = sconst<S>(1.0f);
    float trap = val(r);
    int done = 0;
    for (int i = 0; i < iters; ++i) {
        v3<S> old = z;
        if (fabsf(val(z.x)) > 1.0f) z.x = sconst<S>(val(z.x) > 0.f ? 2.0f : -2.0f) - z.x;
        if (fabsf(val(z.y)) > 1.0f) z.y = sconst<S>(val(z.y) > 0.f ? 2.0f : -2.0f) - z.y;
        if (fabsf(val(z.z)) > 1.0f) z.z = sconst<S>(val(z.z) > 0.f ? 2.0f : -2.0f) - z.z;
        S r2 = dot3(z);
        if (val(r2) < min_r2) {
            z = {s_scale(z.x, min_factor), s_scale(z.y, min_factor), s_scale(z.z, min_factor)};
            dr = s_scale(dr, min_factor);
        } else if (val(r2) < fixed_r2) {
            S f = sconst<S>(fixed_r2) / r2;
            z = {z.x * f, z.y * f, z.z * f};
            dr = dr * f;
        }
        z = {s_scale(z.x, scale), s_scale(z.y, scale), s_scale(z.z, scale)};
        dr = s_scale(dr, abs_scale) + sconst<S>(1.0f);
        z = add(z, p);
        r = orbit_radius(z);
        done = i + 1;
        if (!isfinite(val(r))) {
            z = old;
            r = orbit_radius(z);
            break;
        }
        trap = fminf(trap, val(r));
        if (val(r) > 100.0f) break;
    }
    if (orbit_counter) *orbit_counter += done;
    Leaf<S> out;
    out.d = !(val(dr) > 0.0f) ? r : r / dr;
    out.radius = val(r);
    out.derivative = val(dr);
    out.trap = trap;
    out.iterations = done;
    return out;
}

// Menger leaf (SM:756-784, parametric count/size/divisor, no rotation).
template <class S>
__device__ __forceinline__ S menger_axis_mod(S x, float s) {
    S y = x - sconst<S>(0.5f * s);
    return y - sconst<S>(s * floorf(val(y) / s)) - sconst<S>(0.5f * s);
}
template <class S>
__device__ Leaf<S> leaf_menger(v3<S> p, const float *data, int *orbit_counter) {
    const int iters = int(data[0]);
    const float size = data[1] != 0.0f ? data[1] : 1.0f;
    const float divisor = data[2] != 0.0f ? data[2] : 3.0f;
    const float inv = 1.0f / size;
    v3<S> q = {s_scale(p.x, inv), s_scale(p.y, inv), s_scale(p.z, inv)};
    S sdf = s_max(s_max(s_abs(q.x) - sconst<S>(1.f), s_abs(q.y) - sconst<S>(1.f)), s_abs(q.z) - sconst<S>(1.f));
    float scale = 1.0f;
    for (int i = 0; i < iters; ++i) {
        float s = scale * 2.0f;
        S mx = menger_axis_mod(q.x, s), my = menger_axis_mod(q.y, s), mz = menger_axis_mod(q.z, s);
        S ax = s_abs(mx), ay = s_abs(my), az = s_abs(mz);
        S da = s_max(ax, ay), db = s_max(ay, az), dc = s_max(az, ax);
        S cut = s_min(da, s_min(db, dc)) - sconst<S>(s / 6.0f);
        sdf = s_max(sdf, -cut);
        scale /= divisor;
    }
    if (orbit_counter) *orbit_counter += iters;
    Leaf<S> out;
    out.d = s_scale(sdf, size);
    out.radius = 0.f;
    out.derivative = 0.f;
    out.trap = 0.f;
    out.iterations = iters;
    return out;
}

__device__ __forceinline__ bool is_leaf(uint32_t op) {
    return op == FPT_SDF_OP_MANDELBULB || op == FPT_SDF_OP_MANDELBOX || op == FPT_SDF_OP_MENGER;
}

template <class S>
__device__ __forceinline__ Leaf<S> eval_leaf(uint32_t op, v3<S> p, const float *data, int *oc) {
    if (op == FPT_SDF_OP_MANDELBULB) return leaf_mandelbulb<S>(p, data, oc);
    if (op == FPT_SDF_OP_MANDELBOX) return leaf_mandelbox<S>(p, data, oc);
    return leaf_menger<S>(p, data, oc);
}

// ---------------------------------------------------------------------------
// Typed program interpreter (SM:1065-1069, 1882-2036) with fractal leaves.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float combine(float cur, float cand, uint32_t mode) {
    if (mode == 1u) return fmaxf(cur, cand);
    if (mode == 2u) return fmaxf(cur, -cand);
    return fminf(cur, cand);
}

__device__ __forceinline__ void rot2(float &a, float &b, float angle) {
    float s, c;
    sincosf(angle, &s, &c);
    float na = c * a - s * b;
    b = s * a + c * b;
    a = na;
}

// Apply a transform op to the running point; false when `op` is not one.
__device__ __forceinline__ bool apply_transform(const FptSdfInstruction &op, f3 &p, float &ds) {
    const float *d = op.data;
    switch (op.opcode) {
        case FPT_SDF_OP_ABS: p = vabs(p); return true;
        case FPT_SDF_OP_TRANSLATE: p = p - ld3(d); return true;
        case FPT_SDF_OP_SCALE: {
            float s = fabsf(d[0]) > 1.0e-6f ? d[0] : 1.0f;
            p = p * s;
            ds *= fabsf(s);
            return true;
        }
        case FPT_SDF_OP_ROTATE_X: rot2(p.y, p.z, d[0]); return true;
        case FPT_SDF_OP_ROTATE_Y: rot2(p.x, p.z, d[0]); return true;
        case FPT_SDF_OP_ROTATE_Z: rot2(p.x, p.y, d[0]); return true;
        case FPT_SDF_OP_REPEAT: {
            f3 per = mk(fmaxf(fabsf(d[0]), 1.0e-5f), fmaxf(fabsf(d[1]), 1.0e-5f), fmaxf(fabsf(d[2]), 1.0e-5f));
            p.x -= per.x * floorf(p.x / per.x + 0.5f);
            p.y -= per.y * floorf(p.y / per.y + 0.5f);
            p.z -= per.z * floorf(p.z / per.z + 0.5f);
            return true;
        }
        case FPT_SDF_OP_SORT_DESC: {
            float t;
            if (p.x < p.z) { t = p.x; p.x = p.z; p.z = t; }
            if (p.y < p.z) { t = p.y; p.y = p.z; p.z = t; }
            if (p.x < p.y) { t = p.x; p.x = p.y; p.y = t; }
            return true;
        }
        case FPT_SDF_OP_BOX_FOLD: {
            float l = d[0] != 0.0f ? d[0] : 1.0f;
            p.x = clampf(p.x, -l, l) * 2.0f - p.x;
            p.y = clampf(p.y, -l, l) * 2.0f - p.y;
            p.z = clampf(p.z, -l, l) * 2.0f - p.z;
            return true;
        }
        default: return false;
    }
}

struct Work {  // per-thread work counters (SM:7650-7704 categories)
    int distance_evals;
    int orbit_iterations;
};

// Typed-program evaluation, one opcode at a time.  eval_op<OP>/surf_op<OP>
// are the per-instruction steps of deProgramDistance / deShadingProgram
// (SM:1882-2036) and programSurfaceInterpreted (SM:3266-3369); the
// interpreter below dispatches to them with a switch, and a topology-
// specialised program (runtime-compiled, see specialize.hip) calls the same
// steps with the opcodes fixed at compile time, so both evaluate the same
// float operations in the same order.
struct EvalState {
    f3 p;
    float ds, dist, orbit;
    bool has;
};

template <uint32_t OP>
__device__ __forceinline__ constexpr bool op_is_transform() {
    return OP == FPT_SDF_OP_ABS || OP == FPT_SDF_OP_TRANSLATE || OP == FPT_SDF_OP_SCALE ||
           OP == FPT_SDF_OP_ROTATE_X || OP == FPT_SDF_OP_ROTATE_Y || OP == FPT_SDF_OP_ROTATE_Z ||
           OP == FPT_SDF_OP_REPEAT || OP == FPT_SDF_OP_SORT_DESC || OP == FPT_SDF_OP_BOX_FOLD;
}
template <uint32_t OP>
__device__ __forceinline__ constexpr bool op_is_distance() {
    return OP == FPT_SDF_OP_SPHERE || OP == FPT_SDF_OP_BOX || OP == FPT_SDF_OP_PLANE ||
           OP == FPT_SDF_OP_MANDELBULB || OP == FPT_SDF_OP_MANDELBOX || OP == FPT_SDF_OP_MENGER;
}

template <uint32_t OP, bool ORBIT>
__device__ __forceinline__ void eval_op(const FptSdfInstruction &op, EvalState &s, int *orbit_counter) {
    const float *d = op.data;
    if constexpr (op_is_transform<OP>()) {
        FptSdfInstruction t = op;
        t.opcode = OP;
        (void)apply_transform(t, s.p, s.ds);
    } else if constexpr (OP == FPT_SDF_OP_ORBIT_ADD) {
        if (ORBIT) s.orbit += length(s.p * ld3(d)) * d[3];
    } else if constexpr (op_is_distance<OP>()) {
        const float div = fmaxf(s.ds, 1.0e-6f);
        float cand;
        if constexpr (OP == FPT_SDF_OP_SPHERE) {
            cand = length(s.p) - d[0];
        } else if constexpr (OP == FPT_SDF_OP_BOX) {
            f3 q = vabs(s.p) - vabs(ld3(d));
            cand = fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) + length(vmax0(q));
        } else if constexpr (OP == FPT_SDF_OP_PLANE) {
            cand = dot(s.p, normalize(ld3(d))) + d[3];
        } else {
            Leaf<float> lf = eval_leaf<float>(OP, v3<float>{s.p.x, s.p.y, s.p.z}, d, orbit_counter);
            cand = lf.d;
            if (ORBIT) s.orbit += lf.trap;
        }
        cand = cand / div;
        s.dist = s.has ? combine(s.dist, cand, op.flags) : cand;
        s.has = true;
        if (ORBIT && (OP == FPT_SDF_OP_SPHERE || OP == FPT_SDF_OP_BOX)) s.orbit += fabsf(cand) * d[3];
    }
}

template <bool ORBIT>
__device__ __forceinline__ void eval_dispatch(const FptSdfInstruction &op, EvalState &s, int *oc) {
    switch (op.opcode) {
        case FPT_SDF_OP_ABS: eval_op<FPT_SDF_OP_ABS, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_TRANSLATE: eval_op<FPT_SDF_OP_TRANSLATE, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_SCALE: eval_op<FPT_SDF_OP_SCALE, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_ROTATE_X: eval_op<FPT_SDF_OP_ROTATE_X, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_ROTATE_Y: eval_op<FPT_SDF_OP_ROTATE_Y, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_ROTATE_Z: eval_op<FPT_SDF_OP_ROTATE_Z, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_REPEAT: eval_op<FPT_SDF_OP_REPEAT, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_SORT_DESC: eval_op<FPT_SDF_OP_SORT_DESC, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_BOX_FOLD: eval_op<FPT_SDF_OP_BOX_FOLD, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_SPHERE: eval_op<FPT_SDF_OP_SPHERE, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_BOX: eval_op<FPT_SDF_OP_BOX, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_PLANE: eval_op<FPT_SDF_OP_PLANE, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_MANDELBULB: eval_op<FPT_SDF_OP_MANDELBULB, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_MANDELBOX: eval_op<FPT_SDF_OP_MANDELBOX, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_MENGER: eval_op<FPT_SDF_OP_MENGER, ORBIT>(op, s, oc); break;
        case FPT_SDF_OP_ORBIT_ADD: eval_op<FPT_SDF_OP_ORBIT_ADD, ORBIT>(op, s, oc); break;
        default: break;
    }
}

// deProgramDistance / deShadingProgram: distance (and orbit when asked).
template <bool ORBIT>
__device__ float program_eval(const FptSdfInstruction *ops, uint32_t count, f3 p, float *orbit_out,
                              int *orbit_counter) {
    EvalState s = {p, 1.0f, kInf, 0.0f, false};
    for (uint32_t i = 0; i < count; ++i) eval_dispatch<ORBIT>(ops[i], s, orbit_counter);
    if (ORBIT) *orbit_out = s.orbit;
    return s.dist;
}

struct SurfState {
    f3 p, jx, jy, jz, grad;
    float ds, dist;
    bool has;
};

template <uint32_t OP>
__device__ __forceinline__ void surf_op(const FptSdfInstruction &op, SurfState &s) {
    const float *d = op.data;
    f3 &p = s.p, &jx = s.jx, &jy = s.jy, &jz = s.jz;
    if constexpr (OP == FPT_SDF_OP_ABS) {
        jx = jx * signf(p.x); jy = jy * signf(p.y); jz = jz * signf(p.z);
        p = vabs(p);
    } else if constexpr (OP == FPT_SDF_OP_TRANSLATE) {
        p = p - ld3(d);
    } else if constexpr (OP == FPT_SDF_OP_SCALE) {
        float sc = fabsf(d[0]) > 1.0e-6f ? d[0] : 1.0f;
        p = p * sc; jx = jx * sc; jy = jy * sc; jz = jz * sc;
        s.ds *= fabsf(sc);
    } else if constexpr (OP == FPT_SDF_OP_ROTATE_X) {
        float sn, c; sincosf(d[0], &sn, &c);
        float py = p.y; f3 oj = jy;
        p.y = c * py - sn * p.z; p.z = sn * py + c * p.z;
        jy = oj * c - jz * sn; jz = oj * sn + jz * c;
    } else if constexpr (OP == FPT_SDF_OP_ROTATE_Y) {
        float sn, c; sincosf(d[0], &sn, &c);
        float px = p.x; f3 oj = jx;
        p.x = c * px - sn * p.z; p.z = sn * px + c * p.z;
        jx = oj * c - jz * sn; jz = oj * sn + jz * c;
    } else if constexpr (OP == FPT_SDF_OP_ROTATE_Z) {
        float sn, c; sincosf(d[0], &sn, &c);
        float px = p.x; f3 oj = jx;
        p.x = c * px - sn * p.y; p.y = sn * px + c * p.y;
        jx = oj * c - jy * sn; jy = oj * sn + jy * c;
    } else if constexpr (OP == FPT_SDF_OP_REPEAT) {
        FptSdfInstruction t = op;
        t.opcode = OP;
        float u = 1.f;
        (void)apply_transform(t, p, u);
    } else if constexpr (OP == FPT_SDF_OP_SORT_DESC) {
        float t; f3 tj;
        if (p.x < p.z) { t = p.x; p.x = p.z; p.z = t; tj = jx; jx = jz; jz = tj; }
        if (p.y < p.z) { t = p.y; p.y = p.z; p.z = t; tj = jy; jy = jz; jz = tj; }
        if (p.x < p.y) { t = p.x; p.x = p.y; p.y = t; tj = jx; jx = jy; jy = tj; }
    } else if constexpr (OP == FPT_SDF_OP_BOX_FOLD) {
        float l = d[0] != 0.0f ? d[0] : 1.0f;
        float fx = fabsf(p.x) > l ? -1.f : 1.f, fy = fabsf(p.y) > l ? -1.f : 1.f, fz = fabsf(p.z) > l ? -1.f : 1.f;
        FptSdfInstruction t = op;
        t.opcode = OP;
        float u = 1.f;
        (void)apply_transform(t, p, u);
        jx = jx * fx; jy = jy * fy; jz = jz * fz;
    } else if constexpr (op_is_distance<OP>()) {
        const float scale = fmaxf(s.ds, 1.0e-6f);
        float cand;
        f3 lg;
        if constexpr (OP == FPT_SDF_OP_SPHERE) {
            float radius = length(p);
            cand = (radius - d[0]) / scale;
            lg = radius > 1.0e-8f ? p / radius : mk(0.f, 1.f, 0.f);
        } else if constexpr (OP == FPT_SDF_OP_BOX) {
            f3 q = vabs(p) - vabs(ld3(d));
            f3 o = vmax0(q);
            float ol = length(o);
            cand = (fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) + ol) / scale;
            if (ol > 1.0e-8f) lg = mk(signf(p.x), signf(p.y), signf(p.z)) * o / ol;
            else if (q.x >= q.y && q.x >= q.z) lg = mk(signf(p.x), 0.f, 0.f);
            else if (q.y >= q.z) lg = mk(0.f, signf(p.y), 0.f);
            else lg = mk(0.f, 0.f, signf(p.z));
        } else if constexpr (OP == FPT_SDF_OP_PLANE) {
            lg = normalize(ld3(d));
            cand = (dot(p, lg) + d[3]) / scale;
        } else {
            v3<Df> dp = {{p.x, 1.f, 0.f, 0.f}, {p.y, 0.f, 1.f, 0.f}, {p.z, 0.f, 0.f, 1.f}};
            Leaf<Df> lf = eval_leaf<Df>(OP, dp, d, nullptr);
            cand = lf.d.v / scale;
            lg = mk(lf.d.a, lf.d.b, lf.d.c);
        }
        f3 cg = (jx * lg.x + jy * lg.y + jz * lg.z) / scale;
        bool choose = !s.has || (op.flags == 1u ? cand > s.dist : (op.flags == 2u ? -cand > s.dist : cand < s.dist));
        if (choose) {
            bool neg = op.flags == 2u && s.has;
            s.dist = neg ? -cand : cand;
            s.grad = neg ? -cg : cg;
        }
        s.has = true;
    }
}

__device__ __forceinline__ void surf_dispatch(const FptSdfInstruction &op, SurfState &s) {
    switch (op.opcode) {
        case FPT_SDF_OP_ABS: surf_op<FPT_SDF_OP_ABS>(op, s); break;
        case FPT_SDF_OP_TRANSLATE: surf_op<FPT_SDF_OP_TRANSLATE>(op, s); break;
        case FPT_SDF_OP_SCALE: surf_op<FPT_SDF_OP_SCALE>(op, s); break;
        case FPT_SDF_OP_ROTATE_X: surf_op<FPT_SDF_OP_ROTATE_X>(op, s); break;
        case FPT_SDF_OP_ROTATE_Y: surf_op<FPT_SDF_OP_ROTATE_Y>(op, s); break;
        case FPT_SDF_OP_ROTATE_Z: surf_op<FPT_SDF_OP_ROTATE_Z>(op, s); break;
        case FPT_SDF_OP_REPEAT: surf_op<FPT_SDF_OP_REPEAT>(op, s); break;
        case FPT_SDF_OP_SORT_DESC: surf_op<FPT_SDF_OP_SORT_DESC>(op, s); break;
        case FPT_SDF_OP_BOX_FOLD: surf_op<FPT_SDF_OP_BOX_FOLD>(op, s); break;
        case FPT_SDF_OP_SPHERE: surf_op<FPT_SDF_OP_SPHERE>(op, s); break;
        case FPT_SDF_OP_BOX: surf_op<FPT_SDF_OP_BOX>(op, s); break;
        case FPT_SDF_OP_PLANE: surf_op<FPT_SDF_OP_PLANE>(op, s); break;
        case FPT_SDF_OP_MANDELBULB: surf_op<FPT_SDF_OP_MANDELBULB>(op, s); break;
        case FPT_SDF_OP_MANDELBOX: surf_op<FPT_SDF_OP_MANDELBOX>(op, s); break;
        case FPT_SDF_OP_MENGER: surf_op<FPT_SDF_OP_MENGER>(op, s); break;
        default: break;
    }
}

__device__ __forceinline__ SurfState surf_begin(f3 p) {
    return {p, mk(1.f, 0.f, 0.f), mk(0.f, 1.f, 0.f), mk(0.f, 0.f, 1.f), mk(0.f, 0.f, 0.f), 1.0f, kInf, false};
}

// programSurfaceInterpreted (SM:3266-3369) through fractal leaves.
__device__ float program_surface(const FptSdfInstruction *ops, uint32_t count, f3 p, f3 &grad) {
    SurfState s = surf_begin(p);
    for (uint32_t i = 0; i < count; ++i) surf_dispatch(ops[i], s);
    grad = s.grad;
    return s.dist;
}

#ifdef FPT_TOPOLOGY_SPECIALIZED
// Defined by the runtime-generated source (straight-line eval_op / surf_op
// calls for the program's opcode sequence).
__device__ float topology_distance(f3 p, const FptRenderConfig &cfg, int *orbit_counter);
__device__ float topology_surface(f3 p, const FptRenderConfig &cfg, f3 &grad);
#endif

// ---------------------------------------------------------------------------
// Built-in fractal scenes (SM:730-784).
// ---------------------------------------------------------------------------
__device__ float de_mandelbox_builtin(f3 pos, const FptRenderConfig &cfg, float *orbit, int *oc) {
    const float SCALE = 2.8f + cfg.set_values[2];
    const float MINRAD2 = 0.25f + cfg.set_values[3];
    const float sxyz = SCALE / MINRAD2, sw = fabsf(SCALE) / MINRAD2;
    const float min_rad2 = clampf(MINRAD2, 1.0e-9f, 1.0f);
    const float abs_scale_m1 = fabsf(SCALE - 1.0f);
    // |SCALE|^-9 as an IEEE chain (s^2)^2)^2 * s, so every platform agrees
    // bit for bit (voxel occupancy tests sit on this value).
    const float as1 = fabsf(SCALE), as2 = as1 * as1, as4 = as2 * as2, as8 = as4 * as4;
    const float abs_scale = 1.0f / (as8 * as1);
    const f3 off = mk(cfg.set_values[4], cfg.set_values[5], cfg.set_values[6]);
    f3 p = pos;
    float w = 1.0f;
    float sdf = 0.f, col = 0.f;
    for (int i = 0; i < 12; ++i) {
        p.x = clampf(p.x, -1.f, 1.f) * 2.0f - p.x;
        p.y = clampf(p.y, -1.f, 1.f) * 2.0f - p.y;
        p.z = clampf(p.z, -1.f, 1.f) * 2.0f - p.z;
        float r2 = dot(p, p);
        float k = clampf(fmaxf(min_rad2 / fmaxf(r2, 0.00001f), min_rad2), 0.f, 1.f);
        p = p * k;
        w *= k;
        p = p * sxyz + pos + off;
        w = w * sw + 1.0f;
        sdf = (length(p) - abs_scale_m1) / w - abs_scale;
        col += sdf;
    }
    if (oc) *oc += 12;
    if (orbit) *orbit = col;
    return sdf;
}

__device__ float de_menger_builtin(f3 p, const FptRenderConfig &cfg, int *oc) {
    f3 dd = vabs(p) - mk(1.f, 1.f, 1.f);
    float sdf = fmaxf(fmaxf(dd.x, dd.y), dd.z);
    float scale = fmaxf(cfg.set_values[0], 0.0001f);
    const float divisor = fmaxf(cfg.set_values[1], 0.0001f);
    const float ry = cfg.set_values[2], rp = cfg.set_values[3];
    for (int i = 0; i < 12; ++i) {
        float s = scale * 2.0f;
        float h = s / 2.0f;
        f3 y = mk(p.x - h, p.y - h, p.z - h);
        f3 q = mk(y.x - s * floorf(y.x / s) - h, y.y - s * floorf(y.y / s) - h, y.z - s * floorf(y.z / s) - h);
        float da = fmaxf(fabsf(q.x), fabsf(q.y)), db = fmaxf(fabsf(q.y), fabsf(q.z)), dc = fmaxf(fabsf(q.z), fabsf(q.x));
        sdf = fmaxf(sdf, -(fminf(da, fminf(db, dc)) - s / 6.0f));
        scale /= divisor;
        p = rotate_camera(p, ry, rp);
    }
    if (oc) *oc += 12;
    return sdf;
}

// ---------------------------------------------------------------------------
// Scene dispatch and materials (SM:3418-3615).
// ---------------------------------------------------------------------------
struct Material {
    f3 rgb;
    float roughness, specular, translucency, ior, emission;
};

__device__ __forceinline__ bool is_program(const FptRenderConfig &cfg) {
    return cfg.sdf_id == FPT_SDF_PROGRAM && cfg.sdf_program_count > 0u;
}
__device__ __forceinline__ uint32_t prog_count(const FptRenderConfig &cfg) {
    return cfg.sdf_program_count < 64u ? cfg.sdf_program_count : 64u;
}

// ---------------------------------------------------------------------------
// Direct evaluators of lowered programs (SM:2713-2970): the typed
// structure-of-arrays, canonical affine and flat-union tables the scene front
// end lowers a union of spheres, boxes and planes into.  A program that
// carries them is evaluated from the tables, as the reference does
// (SM:3576-3584 distance, SM:3403-3408 surface).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lowered_cap(uint32_t n) {
    return n < uint32_t(FPT_SDF_FLAT_UNION_MAX_PRIMITIVES) ? n : uint32_t(FPT_SDF_FLAT_UNION_MAX_PRIMITIVES);
}
__device__ __forceinline__ uint32_t typed_soa_count(const FptRenderConfig &cfg) {  // SM:2848-2855
    const FptTypedSoAProgram &t = cfg.sdf_typed_soa;
    return lowered_cap(t.sphere_count) + lowered_cap(t.box_count) + lowered_cap(t.plane_count);
}

// Instance point: the 3x4 row-major affine transform applied to the source.
__device__ __forceinline__ f3 instance_point(f3 s, const float *t) {
    return mk(dot(mk(t[0], t[1], t[2]), s) + t[3], dot(mk(t[4], t[5], t[6]), s) + t[7],
              dot(mk(t[8], t[9], t[10]), s) + t[11]);
}

__device__ __forceinline__ float instance_distance(f3 s, const FptPrimitiveInstance &in) {  // SM:2713-2737
    const f3 p = instance_point(s, in.transform);
    const float div = fmaxf(in.distance_scale, 1.0e-6f);
    if (in.opcode == FPT_SDF_OP_SPHERE) return (length(p) - in.data[0]) / div;
    if (in.opcode == FPT_SDF_OP_BOX) {
        const f3 q = vabs(p) - vabs(ld3(in.data));
        return (fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) + length(vmax0(q))) / div;
    }
    if (in.opcode == FPT_SDF_OP_PLANE) return (dot(p, ld3(in.data)) + in.data[3]) / div;
    return kInf;
}

__device__ float flat_union_distance(f3 s, const FptRenderConfig &cfg) {  // SM:2739-2749
    float d = kInf;
    const uint32_t n = lowered_cap(cfg.sdf_flat_union_count);
    for (uint32_t i = 0; i < n; ++i) d = fminf(d, instance_distance(s, cfg.sdf_flat_union_instances[i]));
    return d;
}

// Canonical CSG list: _pad0 0 union, 1 intersection, 2 subtraction.
__device__ float canonical_distance(f3 s, const FptRenderConfig &cfg) {  // SM:2751-2770
    float d = kInf;
    const uint32_t n = lowered_cap(cfg.sdf_canonical_count);
    for (uint32_t i = 0; i < n; ++i) {
        const FptPrimitiveInstance &in = cfg.sdf_canonical_primitives[i];
        const float c = instance_distance(s, in);
        if (i == 0u) d = c;
        else if (in._pad0 == 1u) d = fmaxf(d, c);
        else if (in._pad0 == 2u) d = fmaxf(d, -c);
        else d = fminf(d, c);
    }
    return d;
}

__device__ float canonical_surface(f3 s, const FptRenderConfig &cfg, f3 &grad) {  // SM:2772-2846
    float d = kInf;
    grad = mk(0.f, 0.f, 0.f);
    const uint32_t n = lowered_cap(cfg.sdf_canonical_count);
    for (uint32_t i = 0; i < n; ++i) {
        const FptPrimitiveInstance &in = cfg.sdf_canonical_primitives[i];
        const float *t = in.transform;
        const f3 r0 = mk(t[0], t[1], t[2]), r1 = mk(t[4], t[5], t[6]), r2 = mk(t[8], t[9], t[10]);
        const f3 p = mk(dot(r0, s) + t[3], dot(r1, s) + t[7], dot(r2, s) + t[11]);
        const float div = fmaxf(in.distance_scale, 1.0e-6f);
        float c = kInf;
        f3 lg = mk(0.f, 1.f, 0.f);
        if (in.opcode == FPT_SDF_OP_SPHERE) {
            const float r = length(p);
            c = (r - in.data[0]) / div;
            if (r > 1.0e-8f) lg = p / r;
        } else if (in.opcode == FPT_SDF_OP_BOX) {
            const f3 q = vabs(p) - vabs(ld3(in.data));
            const f3 o = vmax0(q);
            const float ol = length(o);
            c = (fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) + ol) / div;
            if (ol > 1.0e-8f) lg = mk(signf(p.x), signf(p.y), signf(p.z)) * o / ol;
            else if (q.x >= q.y && q.x >= q.z) lg = mk(signf(p.x), 0.f, 0.f);
            else if (q.y >= q.z) lg = mk(0.f, signf(p.y), 0.f);
            else lg = mk(0.f, 0.f, signf(p.z));
        } else if (in.opcode == FPT_SDF_OP_PLANE) {
            lg = ld3(in.data);
            c = (dot(p, lg) + in.data[3]) / div;
        }
        const f3 cg = (r0 * lg.x + r1 * lg.y + r2 * lg.z) / div;
        const bool choose = i == 0u || (in._pad0 == 1u ? c > d : (in._pad0 == 2u ? -c > d : c < d));
        if (choose) {
            const bool neg = in._pad0 == 2u && i > 0u;
            d = neg ? -c : c;
            grad = neg ? -cg : cg;
        }
    }
    return d;
}

__device__ float typed_soa_distance(f3 s, const FptRenderConfig &cfg) {  // SM:2857-2896
    const FptTypedSoAProgram &t = cfg.sdf_typed_soa;
    float d = kInf;
    for (uint32_t i = 0, n = lowered_cap(t.sphere_count); i < n; ++i)
        d = fminf(d, length(s - mk(t.sphere_x[i], t.sphere_y[i], t.sphere_z[i])) - t.sphere_radius[i]);
    for (uint32_t i = 0, n = lowered_cap(t.box_count); i < n; ++i) {
        const f3 q = vabs(s - mk(t.box_x[i], t.box_y[i], t.box_z[i])) - mk(t.box_half_x[i], t.box_half_y[i], t.box_half_z[i]);
        d = fminf(d, fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) + length(vmax0(q)));
    }
    for (uint32_t i = 0, n = lowered_cap(t.plane_count); i < n; ++i) {
        const f3 nrm = normalize(mk(t.plane_x[i], t.plane_y[i], t.plane_z[i]));
        const f3 c = mk(t.plane_center_x[i], t.plane_center_y[i], t.plane_center_z[i]);
        d = fminf(d, dot(s - c, nrm) + t.plane_offset[i]);
    }
    return d;
}

// Union winner with its gradient; ties go to the earlier source instruction.
__device__ float typed_soa_surface(f3 s, const FptRenderConfig &cfg, f3 &grad) {  // SM:2898-2970
    const FptTypedSoAProgram &t = cfg.sdf_typed_soa;
    float d = kInf;
    grad = mk(0.f, 0.f, 0.f);
    uint32_t winner = 0xffffffffu;
    for (uint32_t i = 0, n = lowered_cap(t.sphere_count); i < n; ++i) {
        const f3 l = s - mk(t.sphere_x[i], t.sphere_y[i], t.sphere_z[i]);
        const float r = length(l), c = r - t.sphere_radius[i];
        if (c < d || (c == d && t.sphere_source[i] < winner)) {
            d = c;
            grad = r > 1.0e-8f ? l / r : mk(0.f, 1.f, 0.f);
            winner = t.sphere_source[i];
        }
    }
    for (uint32_t i = 0, n = lowered_cap(t.box_count); i < n; ++i) {
        const f3 l = s - mk(t.box_x[i], t.box_y[i], t.box_z[i]);
        const f3 q = vabs(l) - mk(t.box_half_x[i], t.box_half_y[i], t.box_half_z[i]);
        const f3 o = vmax0(q);
        const float ol = length(o);
        const float c = fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f) + ol;
        if (c < d || (c == d && t.box_source[i] < winner)) {
            d = c;
            if (ol > 1.0e-8f) grad = mk(signf(l.x), signf(l.y), signf(l.z)) * o / ol;
            else if (q.x >= q.y && q.x >= q.z) grad = mk(signf(l.x), 0.f, 0.f);
            else if (q.y >= q.z) grad = mk(0.f, signf(l.y), 0.f);
            else grad = mk(0.f, 0.f, signf(l.z));
            winner = t.box_source[i];
        }
    }
    for (uint32_t i = 0, n = lowered_cap(t.plane_count); i < n; ++i) {
        const f3 nrm = normalize(mk(t.plane_x[i], t.plane_y[i], t.plane_z[i]));
        const f3 c0 = mk(t.plane_center_x[i], t.plane_center_y[i], t.plane_center_z[i]);
        const float c = dot(s - c0, nrm) + t.plane_offset[i];
        if (c < d || (c == d && t.plane_source[i] < winner)) {
            d = c;
            grad = nrm;
            winner = t.plane_source[i];
        }
    }
    return d;
}

__device__ __forceinline__ float distance_sdf(f3 p, const FptRenderConfig &cfg, Work &w) {
    w.distance_evals++;
#ifdef FPT_TOPOLOGY_SPECIALIZED
    if (is_program(cfg)) return topology_distance(p, cfg, &w.orbit_iterations);
#endif
    if (is_program(cfg)) return program_eval<false>(cfg.sdf_program, prog_count(cfg), p, nullptr, &w.orbit_iterations);
    if (cfg.sdf_id == FPT_SDF_MANDELBOX_FRACTAL) return de_mandelbox_builtin(p, cfg, nullptr, &w.orbit_iterations);
    if (cfg.sdf_id == FPT_SDF_MENGER_SPONGE) return de_menger_builtin(p, cfg, &w.orbit_iterations);
    return 1000.0f;
}

__device__ __forceinline__ float map_sdf(f3 p, const FptRenderConfig &cfg, Work &w) {  // SM:3612-3615
    float d = distance_sdf(p, cfg, w);
    return mixf(d, d / 2.0f, cfg.render[6]);
}

// The scene's distance and map with the lowered-program evaluators (the
// normal is normal_scene, after normal_at):
// a program that carries typed-SoA, canonical or flat-union tables is
// evaluated from them (SM:3576-3584, 3403-3408); everything else is
// distance_sdf / map_sdf / normal_at.  (A runtime-specialised build evaluates
// its topology instead, as the reference's does.)  The render kernels have
// instances with and without the tables' code: only configs that carry
// tables pay for it.
__device__ __forceinline__ bool has_lowered(const FptRenderConfig &cfg) {
    return is_program(cfg) && (typed_soa_count(cfg) > 0u || cfg.sdf_canonical_count > 0u || cfg.sdf_flat_union_count > 0u);
}
__device__ __forceinline__ float distance_scene(f3 p, const FptRenderConfig &cfg, Work &w) {
#ifndef FPT_TOPOLOGY_SPECIALIZED
    if (has_lowered(cfg)) {
        w.distance_evals++;
        if (typed_soa_count(cfg) > 0u) return typed_soa_distance(p, cfg);
        if (cfg.sdf_canonical_count > 0u) return canonical_distance(p, cfg);
        return flat_union_distance(p, cfg);
    }
#endif
    return distance_sdf(p, cfg, w);
}
__device__ __forceinline__ float map_scene(f3 p, const FptRenderConfig &cfg, Work &w) {  // SM:3612-3615
    const float d = distance_scene(p, cfg, w);
    return mixf(d, d / 2.0f, cfg.render[6]);
}
__device__ __forceinline__ f3 hsv2rgb(f3 c) {  // SM:534-539
    f3 p = mk(fabsf(fractf(c.x + 1.0f) * 6.0f - 3.0f), fabsf(fractf(c.x + 2.0f / 3.0f) * 6.0f - 3.0f),
              fabsf(fractf(c.x + 1.0f / 3.0f) * 6.0f - 3.0f));
    f3 q = mk(clampf(p.x - 1.f, 0.f, 1.f), clampf(p.y - 1.f, 0.f, 1.f), clampf(p.z - 1.f, 0.f, 1.f));
    return vmix(mk(1.f, 1.f, 1.f), q, c.y) * c.z;
}

__device__ f3 program_gradient(float value, const FptRenderConfig &cfg) {  // SM:3418-3432
    uint32_t count = cfg.gradient_count < 16u ? cfg.gradient_count : 16u;
    if (count == 0u) return ld3(cfg.program_material);
    if (count == 1u) return ld3(&cfg.gradient_stops[0][1]);
    float t = fractf(value);
    const float *prev = cfg.gradient_stops[0];
    for (uint32_t i = 1; i < count; ++i) {
        const float *cur = cfg.gradient_stops[i];
        if (t <= cur[0]) {
            float f = clampf((t - prev[0]) / fmaxf(cur[0] - prev[0], 1.0e-6f), 0.f, 1.f);
            return vmix(ld3(prev + 1), ld3(cur + 1), f);
        }
        prev = cur;
    }
    return ld3(prev + 1);
}

// The reference's material is a function of one scalar orbit value
// (SM:3434-3483): split into orbit -> shading parameter -> material so the
// voxel builder can index materials by a quantised parameter.
__device__ float material_orbit(f3 p, const FptRenderConfig &cfg, Work &w) {
    if (is_program(cfg)) {
        float orbit = 0.f;
        // Only modes 1 and 2 colour by orbit; mode 0 skips the shading program.
        if (cfg.material_mode == 1u || cfg.material_mode == 2u) {
            uint32_t n = cfg.sdf_shading_program_count < 64u ? cfg.sdf_shading_program_count : 64u;
            program_eval<true>(cfg.sdf_shading_program, n, p, &orbit, &w.orbit_iterations);
        }
        return orbit;
    }
    float orbit = 0.f;
    if (cfg.sdf_id == FPT_SDF_MANDELBOX_FRACTAL) de_mandelbox_builtin(p, cfg, &orbit, &w.orbit_iterations);
    if (cfg.sdf_id == FPT_SDF_MENGER_SPONGE) orbit = 1.0f;
    return orbit;
}

__device__ __forceinline__ float material_param(float orbit, const FptRenderConfig &cfg) {
    if (!is_program(cfg) && cfg.fractal_style_mode != 0u) return orbit * cfg.fractal_style[1] + cfg.fractal_style[0];
    return orbit * (1.0f + cfg.set_values[1]) + cfg.set_values[0];
}

// True when the material does not depend on the parameter.
__device__ __forceinline__ bool material_is_constant(const FptRenderConfig &cfg) {
    if (is_program(cfg)) return cfg.material_mode != 1u && cfg.material_mode != 2u;
    if (cfg.fractal_style_mode == 2u) return true;
    return cfg.fractal_style_mode == 0u && cfg.sdf_id == FPT_SDF_MENGER_SPONGE;
}

__device__ Material material_at_param(float o, const FptRenderConfig &cfg) {
    Material m;
    m.rgb = mk(1.f, 1.f, 1.f);
    m.roughness = 1.f; m.specular = 0.f; m.translucency = 0.f; m.ior = 1.5f; m.emission = 0.f;
    if (is_program(cfg)) {  // SM:3434-3449
        if (cfg.material_mode == 1u) m.rgb = hsv2rgb(mk(o, 0.5f, 1.0f));
        else if (cfg.material_mode == 2u) m.rgb = program_gradient(o, cfg);
        else m.rgb = ld3(cfg.program_material);
        m.roughness = cfg.program_material[3];
        m.specular = cfg.program_material[4];
        m.translucency = cfg.program_material[5];
        m.ior = cfg.program_material[6];
        m.emission = cfg.program_material[7];
        return m;
    }
    if (cfg.fractal_style_mode != 0u) {  // SM:3451-3468
        if (cfg.fractal_style_mode == 1u) m.rgb = hsv2rgb(mk(o, cfg.fractal_style[2], cfg.fractal_style[3]));
        else if (cfg.fractal_style_mode == 2u) m.rgb = ld3(&cfg.fractal_style[8]) * cfg.fractal_style[3];
        else m.rgb = program_gradient(o, cfg) * cfg.fractal_style[3];
        m.rgb = vmix(m.rgb, ld3(&cfg.fractal_style[8]), clampf(cfg.fractal_style[7], 0.f, 1.f));
        m.roughness = cfg.fractal_style[4];
        m.specular = cfg.fractal_style[5];
        m.emission = cfg.fractal_style[6];
        return m;
    }
    if (cfg.sdf_id == FPT_SDF_MENGER_SPONGE) return m;  // SM:3469-3472
    m.rgb = hsv2rgb(mk(o, 0.5f, 1.0f));
    return m;
}

__device__ Material user_material(f3 p, const FptRenderConfig &cfg, Work &w) {  // SM:3485-3562
    return material_at_param(material_param(material_orbit(p, cfg, w), cfg), cfg);
}

// march (SM:3753-3783), non-Mandelbulber branch.  LOWERED: the distance
// reads lowered tables when the config carries them (map_scene); without it
// the march is map_sdf's, for kernel instances whose configs carry none.
template <bool LOWERED = true>
__device__ f3 march(f3 dr, f3 rp, int ni, float min_dist, float lod_falloff, const FptRenderConfig &cfg,
                    Work &w, int &steps) {
    const f3 cam = rp;
    const int max_iter = ni < 360 ? ni : 360;
    const bool check_t = cfg.sdf_id == FPT_SDF_PROGRAM && cfg.program_material[5] > 0.0f;
    const float adapt = cfg.render[5];
    const float maxd = cfg.render[4];
    steps = 0;
    for (int i = 0; i < max_iter; ++i) {
        float d = LOWERED ? map_scene(rp, cfg, w) : map_sdf(rp, cfg, w);
        ++steps;
        if (!isfinite(d)) break;
        float o = fabsf(d) * 0.99f;
        rp = rp + dr * o;
        float lod = min_dist;
        if (lod_falloff > 0.00001f) {
            f3 df = cam - rp;
            float l2 = mixf(lod, 0.1f, dot(df, df) * lod_falloff);
            lod = mixf(min_dist, l2, adapt);
        }
        if (check_t && user_material(rp, cfg, w).translucency > 0.0f) lod = 0.0001f;
        if (o < lod) break;
        if (maxd < o) break;
    }
    return rp;
}

// estimateFocusDistance (SM:5869-5878): the centre ray's hit distance, or 5
// when it misses, is degenerate or passes 0.99 max_distance.
__device__ float focus_estimate(const FptRenderConfig &cfg, Work &w) {
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    const f3 dr = rotate_camera_roll(normalize(mk(0.f, 0.f, focal)), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1],
                                     cfg.camera_roll);
    const f3 cam = ld3(cfg.camera_position);
    int steps = 0;
    const f3 hit = march(dr, cam, 120, 0.001f, 0.001f, cfg, w, steps);
    const float d = length(hit - cam);
    if (!isfinite(d) || d <= 0.001f || d > cfg.render[4] * 0.99f) return 5.0f;
    return d;
}

// Thin-lens depth of field of the camera ray (SM:5950-5956): the ray keeps
// its point at the focus distance and starts from a lens offset of radius
// camera_dof.  Every render entry point runs the focus prepass before its
// kernels (focus_prepass_kernel), so a config with depth of field always
// carries a positive focus distance here.
__device__ __forceinline__ void apply_dof(UV xy, float frame, const FptRenderConfig &cfg, f3 &rp, f3 &dr) {
    if (!(cfg.camera_dof > 0.0f)) return;
    const f3 fp = rp + dr * cfg.focus_distance;
    float lx, ly;
    random_point(cfg.camera_dof, xy, frame, lx, ly);
    rp = rp + rotate_camera_roll(mk(lx, ly, 0.f), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1], cfg.camera_roll);
    dr = normalize(fp - rp);
}

__device__ __forceinline__ f3 normalize_fd(f3 g) {  // SM:3863-3869
    float m = fmaxf(fabsf(g.x), fmaxf(fabsf(g.y), fabsf(g.z)));
    if (!(m > 0.0f) || !isfinite(m)) return mk(0.f, 1.f, 0.f);
    return normalize(g / m);
}

// normalAt (SM:3871-3971), non-Mandelbulber branches.
__device__ f3 normal_at(f3 p, const FptRenderConfig &cfg, Work &w) {
    const float e = fmaxf(cfg.render[2], 0.0002f);
    if ((cfg.sdf_normal_mode == 0u || cfg.sdf_normal_mode == 2u) && is_program(cfg)) {
        f3 g;
#ifdef FPT_TOPOLOGY_SPECIALIZED
        topology_surface(p, cfg, g);
#else
        program_surface(cfg.sdf_program, prog_count(cfg), p, g);
#endif
        if (dot(g, g) >= 1.0e-12f && finite3(g)) return normalize(g);
    }
    if (cfg.sdf_normal_mode == 1u) {
        float te = e * 0.57735026919f;
        f3 k0 = mk(1.f, -1.f, -1.f), k1 = mk(-1.f, -1.f, 1.f), k2 = mk(-1.f, 1.f, -1.f), k3 = mk(1.f, 1.f, 1.f);
        float d0 = map_sdf(p + k0 * te, cfg, w), d1 = map_sdf(p + k1 * te, cfg, w);
        float d2 = map_sdf(p + k2 * te, cfg, w), d3 = map_sdf(p + k3 * te, cfg, w);
        return normalize_fd(k0 * d0 + k1 * d1 + k2 * d2 + k3 * d3);
    }
    f3 ex = mk(e, 0.f, 0.f), ey = mk(0.f, e, 0.f), ez = mk(0.f, 0.f, e);
    f3 n = mk(map_sdf(p + ex, cfg, w) - map_sdf(p - ex, cfg, w), map_sdf(p + ey, cfg, w) - map_sdf(p - ey, cfg, w),
              map_sdf(p + ez, cfg, w) - map_sdf(p - ez, cfg, w));
    return normalize_fd(n);
}

__device__ f3 normal_scene(f3 p, const FptRenderConfig &cfg, Work &w) {  // SM:3871-3971
#ifndef FPT_TOPOLOGY_SPECIALIZED
    if (has_lowered(cfg)) {
        const float e = fmaxf(cfg.render[2], 0.0002f);
        if (cfg.sdf_normal_mode == 0u || cfg.sdf_normal_mode == 2u) {
            // programSurface: typed SoA, canonical, else the interpreted
            // program (a flat union has no surface form)
            f3 g;
            if (typed_soa_count(cfg) > 0u) typed_soa_surface(p, cfg, g);
            else if (cfg.sdf_canonical_count > 0u) canonical_surface(p, cfg, g);
            else program_surface(cfg.sdf_program, prog_count(cfg), p, g);
            if (dot(g, g) >= 1.0e-12f && finite3(g)) return normalize(g);
        }
        if (cfg.sdf_normal_mode == 1u) {
            const float te = e * 0.57735026919f;
            const f3 k0 = mk(1.f, -1.f, -1.f), k1 = mk(-1.f, -1.f, 1.f), k2 = mk(-1.f, 1.f, -1.f), k3 = mk(1.f, 1.f, 1.f);
            const float d0 = map_scene(p + k0 * te, cfg, w), d1 = map_scene(p + k1 * te, cfg, w);
            const float d2 = map_scene(p + k2 * te, cfg, w), d3 = map_scene(p + k3 * te, cfg, w);
            return normalize_fd(k0 * d0 + k1 * d1 + k2 * d2 + k3 * d3);
        }
        const f3 ex = mk(e, 0.f, 0.f), ey = mk(0.f, e, 0.f), ez = mk(0.f, 0.f, e);
        return normalize_fd(mk(map_scene(p + ex, cfg, w) - map_scene(p - ex, cfg, w),
                               map_scene(p + ey, cfg, w) - map_scene(p - ey, cfg, w),
                               map_scene(p + ez, cfg, w) - map_scene(p - ez, cfg, w)));
    }
#endif
    return normal_at(p, cfg, w);
}

// ---------------------------------------------------------------------------
// Lighting (SM:5764-5861).
// ---------------------------------------------------------------------------
__device__ f3 sky(f3 dr, f3 sun) {  // SM:5764-5787
    const float T = 2.0f;
    float A = 0.1787f * T - 1.4630f, B = -0.3554f * T + 0.4275f, C = -0.0227f * T + 5.3251f;
    float D = 0.1206f * T - 2.5771f, E = -0.0670f * T + 0.3703f;
    float theta = acosf(clampf(dr.y, -1.f, 1.f));
    float gamma = acosf(clampf(dot(dr, sun), -1.f, 1.f));
    float cg = cosf(gamma);
    float sk = (1.0f + A * expf(B / fmaxf(0.1f, cosf(theta)))) * (1.0f + C * expf(D * gamma) + E * cg * cg);
    f3 col = vmix(mk(0.004f, 0.048f, 0.253f), mk(0.8f, 0.9f, 1.0f), sk);
    float hd = (dr.y + 1.0f) / 2.0f, hs = (sun.y + 1.0f) / 2.0f;
    col = vmix(mk(1.0f, 0.85f, 0.53f) / 4.0f, col * 1.2f, hd) * hs;
    col = (col - mk(0.5f, 0.5f, 0.5f)) * 1.2f + mk(0.5f, 0.5f, 0.5f);
    float disk = fmaxf((dot(dr, sun) - 1.0f) / (1.0f - cosf(0.05f)) + 1.0f, 0.0f);
    disk *= powf(hd, 10.0f);
    return col * 0.8f + mk(disk, disk, disk) * 100.0f;
}

// One channel of an HDRI LUT texel (SM:5796-5801): FP16, row-major, RGB.
__device__ __forceinline__ float hdri_texel(const FptRenderConfig &cfg, uint32_t x, uint32_t y, uint32_t c, uint32_t w,
                                            uint32_t h) {
    const uint32_t i = ((y < h - 1u ? y : h - 1u) * w + (x < w - 1u ? x : w - 1u)) * 3u + c;
    return float(__builtin_bit_cast(_Float16, cfg.hdri_lut[i]));
}

// sampleHdriLut (SM:5803-5821): bilinear lookup of the equirectangular LUT
// (u = azimuth atan2(z, x), wrapping; v = polar angle from +y, clamped).
// fract follows the shading language's definition, which never returns 1.
__device__ f3 sample_hdri(f3 direction, const FptRenderConfig &cfg) {
    const uint32_t w = cfg.hdri_width < 1u ? 1u : (cfg.hdri_width > 32u ? 32u : cfg.hdri_width);
    const uint32_t h = cfg.hdri_height < 1u ? 1u : (cfg.hdri_height > 16u ? 16u : cfg.hdri_height);
    auto fr = [](float x) { return fminf(x - floorf(x), 0x1.fffffep-1f); };
    const f3 d = normalize(direction);
    const float u = fr(atan2f(d.z, d.x) / (2.0f * kPi) + 0.5f);
    const float v = clampf(acosf(clampf(d.y, -1.0f, 1.0f)) / kPi, 0.0f, 1.0f);
    const float x = u * float(w) - 0.5f;
    const float y = v * float(h - 1u);
    const int x0i = int(floorf(x));
    const uint32_t x0 = uint32_t((x0i % int(w) + int(w)) % int(w));
    const uint32_t x1 = (x0 + 1u) % w;
    const int y0i = int(floorf(y));
    const uint32_t y0 = uint32_t(y0i < 0 ? 0 : (y0i > int(h - 1u) ? int(h - 1u) : y0i));
    const uint32_t y1 = y0 + 1u < h - 1u ? y0 + 1u : h - 1u;
    const float fx = fr(x), fy = fr(y);
    auto texel = [&](uint32_t tx, uint32_t ty) {
        return mk(hdri_texel(cfg, tx, ty, 0u, w, h), hdri_texel(cfg, tx, ty, 1u, w, h), hdri_texel(cfg, tx, ty, 2u, w, h));
    };
    return vmix(vmix(texel(x0, y0), texel(x1, y0), fx), vmix(texel(x0, y1), texel(x1, y1), fx), fy);
}

__device__ f3 environment(f3 view, const FptRenderConfig &cfg) {  // SM:5823-5844
    if (cfg.world[4] < 0.00001f) return mk(0.f, 0.f, 0.f);
    f3 ld = rotate_camera(mk(0.f, 0.f, 1.f), cfg.world[2] * kPi / 180.0f, cfg.world[3] * kPi / 180.0f);
    f3 env;
    if (cfg.world[0] == 0.0f) {
        float ls = fmaxf(cfg.world[1], 0.0001f);
        float l = fmaxf((dot(view, ld) - 1.0f) / (1.0f - cosf(ls)) + 1.0f, 0.0f) / ls * 3.0f;
        env = mk(l, l, l);
    } else if (cfg.world[0] == 1.0f) {
        env = sky(view, ld);
    } else if (cfg.world[0] == 2.0f && cfg.hdri_enabled != 0u && cfg.hdri_width > 0u) {
        env = sample_hdri(rotate_camera(view, cfg.world[2] * kPi / 180.0f, 0.0f), cfg);
    } else if (cfg.world[0] == 3.0f) {
        env = ld3(cfg.world_one_color);
    } else {
        env = mk(1.f, 1.f, 1.f);
    }
    float contrast = cfg.world[0] == 3.0f ? 1.0f : cfg.world[5];
    return vmax0(env * cfg.world[4] - mk(0.5f, 0.5f, 0.5f) * contrast + mk(0.5f, 0.5f, 0.5f));
}

// ---------------------------------------------------------------------------
// Mandelbulber native field: Kaleidoscopic IFS, formula 10
// (Shaders.metal:786-949; CPU form src/mandelbulber/orbit.rs:699-750 with
// distances 0, intensities 1, no per-plane rotation, Ifs finalizer
// orbit.rs:662).  Parameters: set_values[0] world scale, [1] iterations,
// [2] bailout, [3] IFS scale, [4] abs mask, [5] plane mask, [6..8] offset,
// [9..11] rotation, [12 + 3k ..] plane k direction (k < 9); vset_values
// [115..119, 130] the screen iteration budget, [120..128] the global
// offset / rotation / repeat.
// ---------------------------------------------------------------------------
// Rz(r.x) Ry(r.y) Rx(r.z) (Mandelbulber SetRotation3), applied x first.
__device__ __forceinline__ f3 mb_rotation3(f3 v, f3 r) {
    float s, c;
    sincosf(r.z, &s, &c);
    v = mk(v.x, v.y * c - v.z * s, v.y * s + v.z * c);
    sincosf(r.y, &s, &c);
    v = mk(v.x * c + v.z * s, v.y, -v.x * s + v.z * c);
    sincosf(r.x, &s, &c);
    return mk(v.x * c - v.y * s, v.x * s + v.y * c, v.z);
}
// Rz(r.z) Ry(r.y) Rx(r.x) (SetRotation2).
__device__ __forceinline__ f3 mb_rotation2(f3 v, f3 r) {
    return mb_rotation3(v, mk(r.z, r.y, r.x));
}
__device__ __forceinline__ float mb_repeat(float v, float period) {
    if (!(period > 0.0f)) return v;
    return v - period * floorf((v + 0.5f * period) / period);
}

__device__ __forceinline__ f3 mb_global_point(f3 p, const FptRenderConfig &cfg) {  // SM:827-843
    const float ws = fmaxf(cfg.set_values[0], 1.0f);
    const float *v = cfg.vset_values;
    f3 q = mk(p.x, p.z, p.y) / ws - mk(v[120], v[121], v[122]);
    q = mb_rotation2(q, mk(v[123], v[124], v[125]));
    q = mk(mb_repeat(q.x, v[126]), mb_repeat(q.y, v[127]), mb_repeat(q.z, v[128]));
    return mk(q.x, q.z, q.y) * ws;
}

__device__ __forceinline__ int mb_iteration_budget(f3 world, const FptRenderConfig &cfg, int mult) {  // SM:863-884
    const float *v = cfg.vset_values;
    int configured = int(cfg.set_values[1]) * (mult > 1 ? mult : 1);
    configured = configured < 1 ? 1 : (configured > 4096 ? 4096 : configured);
    const float rate = v[130];
    if (!(rate > 0.0f) || !isfinite(rate)) return configured;
    float fp = v[115] > 0.5f ? length(ld3(cfg.camera_position) - world) * v[116] : v[117];
    fp = clampf(fp, v[118], v[119]);
    const float minimum = fmaxf(v[118], 1.0e-12f);
    const float octaves = fmaxf(log2f(fmaxf(fp / minimum, 1.0f)), 0.0f);
    const int reduction = int(floorf(octaves * rate));
    int floor_budget = int(ceilf(float(configured) * 0.75f));
    floor_budget = floor_budget > 1 ? floor_budget : 1;
    return configured - reduction > floor_budget ? configured - reduction : floor_budget;
}

// {distance, radius, derivative, +-iterations (negative = escaped)}.
__device__ float4 kifs_field_sample(f3 p, const FptRenderConfig &cfg, int mult) {  // SM:897-949
    const float *sv = cfg.set_values;
    const int budget = mb_iteration_budget(p, cfg, mult);
    p = mb_global_point(p, cfg);
    const float ws = fmaxf(sv[0], 1.0f);
    const f3 sc = p / ws;
    f3 z = mk(sc.x, sc.z, sc.y);
    float der = 1.0f, r = length(z);
    int done = 0;
    bool escaped = false;
    const float bailout = fmaxf(sv[2], 1.0f), k = fmaxf(sv[3], 1.000001f);
    const uint32_t abs_mask = uint32_t(rintf(sv[4])), plane_mask = uint32_t(rintf(sv[5]));
    const f3 off = mk(sv[6], sv[7], sv[8]), rot = mk(sv[9], sv[10], sv[11]);
    for (int it = 0; it < budget; ++it) {
        if (abs_mask & 1u) z.x = fabsf(z.x);
        if (abs_mask & 2u) z.y = fabsf(z.y);
        if (abs_mask & 4u) z.z = fabsf(z.z);
        for (uint32_t pl = 0; pl < 9u; ++pl) {
            if (!(plane_mask & (1u << pl))) continue;
            const f3 dir = mk(sv[12 + 3 * pl], sv[13 + 3 * pl], sv[14 + 3 * pl]);
            const float proj = dot(z, dir);
            if (proj < 0.0f) z = z - dir * (2.0f * proj);
        }
        z = mb_rotation3(z - off, rot) + off;
        z = z * k - off * (k - 1.0f);
        der *= fabsf(k);
        r = length(z);
        done = it + 1;
        if (r > bailout) {
            escaped = true;
            break;
        }
    }
    const float d = (r - 2.0f) / fmaxf(fabsf(der), 1.0e-30f) * ws;
    return make_float4(d, r, der, escaped ? -float(done) : float(done));
}

// ---------------------------------------------------------------------------
// Rendering imported Mandelbulber scenes (sdf_id MANDELBULBER): the
// reference's Mandelbulber branches of the camera (SM:473-511), the distance
// (SM:975-995), the march (SM:3627-3751) and the normal (SM:3871-3960).
// Parameters: vset_values[107] projection (0 perspective, 1 fish-eye,
// 2 equirectangular, 3 fish-eye cut at 180 degrees), [108..113] step
// clamps and DE factor, [114] normal step per threshold, [115] threshold
// mode (0 fixed [117], 1 distance-scaled by [116], 2 iteration-limit
// interior), [118..119] threshold clamp, [129] detail level.
//
// Only the one-path-per-lane kernel and the per-ray hooks call these, through
// the *_any entry points below; the persistent kernel's march, material and
// normal stay as they are.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float mb_threshold(f3 p, const FptRenderConfig &cfg) {  // SM:3634-3640
    const float *v = cfg.vset_values;
    const float t = v[115] > 0.5f ? length(ld3(cfg.camera_position) - p) * v[116] : v[117];
    return clampf(t, v[118], v[119]);
}

// deMandelbulber: iteration-limit interior in mode 2, orbit = |iterations| /
// configured iterations.
__device__ float de_kifs(f3 p, const FptRenderConfig &cfg, float *orbit, Work &w) {  // SM:975-995
    const float4 s = kifs_field_sample(p, cfg, 1);
    w.orbit_iterations += int(fabsf(s.w));
    float d = s.x;
    const float *v = cfg.vset_values;
    if (v[115] > 1.5f) {
        const float thr = clampf(length(ld3(cfg.camera_position) - p) * v[116], v[118], v[119]);
        if (s.w > 0.0f) d = 0.0f;
        else if (d < thr) d = thr * 1.01f;
    }
    if (orbit) *orbit = fabsf(s.w) / fmaxf(cfg.set_values[1], 1.0f);
    return d;
}

__device__ __forceinline__ float map_kifs(f3 p, const FptRenderConfig &cfg, Work &w) {  // SM:3612-3615
    w.distance_evals++;
    const float d = de_kifs(p, cfg, nullptr, w);
    return mixf(d, d / 2.0f, cfg.render[6]);
}

__device__ __forceinline__ bool mb_projection_visible(float x, float y, const FptRenderConfig &cfg) {  // SM:473-480
    const float proj = cfg.vset_values[107];
    if (proj < 2.5f || proj > 3.5f) return true;
    const float fov = cfg.camera_fov * kPi / 180.0f;
    return sqrtf(x * x + y * y) <= 0.5f * kPi / fmaxf(fov, 1.0e-6f);
}

__device__ f3 mb_camera_ray(float x, float y, const FptRenderConfig &cfg) {  // SM:482-511
    const float proj = cfg.vset_values[107];
    const float fov = cfg.camera_fov * kPi / 180.0f;
    f3 d;
    if ((proj > 0.5f && proj < 1.5f) || (proj > 2.5f && proj < 3.5f)) {
        const float r = sqrtf(x * x + y * y);
        if (r <= 1.0e-8f) {
            d = mk(0.f, 0.f, 1.f);
        } else {
            const float k = sinf(r * fov) / r;
            d = mk(x * k, y * k, cosf(r * fov));
        }
    } else if (proj > 1.5f && proj < 2.5f) {
        const float aspect = float(cfg.width) / float(cfg.height);
        const float ax = x * (2.0f / aspect) * (0.5f * fov), ay = y * (0.5f * fov);
        d = mk(sinf(ax) * cosf(ay), sinf(ay), cosf(ax) * cosf(ay));
    } else {
        d = normalize(mk(x, y, 1.0f / tanf(0.5f * fov)));
    }
    return rotate_camera_roll(normalize(d), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1], cfg.camera_roll);
}

__device__ __forceinline__ float mb_step(float d, float thr, const FptRenderConfig &cfg) {  // SM:3642-3676
    const float *v = cfg.vset_values;
    float step = fmaxf(d - 0.5f * thr, 0.0f) * v[113];
    if (v[108] > 0.5f) {
        step = clampf(step, v[109], v[110]);
        if (thr > v[109]) step = clampf(step, v[111] * thr, v[112] * thr);
    } else {
        step = fminf(step, v[110]);
    }
    return step;
}

__device__ __forceinline__ bool same3(f3 a, f3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// marchMandelbulber: march until d < threshold (the threshold grows with
// camera distance), stopping where FP32 no longer moves the point, then up
// to 30 half-steps bracketing the band (threshold * (1 - 0.001 detail),
// threshold).  steps counts the march iterations, not the refinement;
// found (optional) reports whether the band was reached.
__device__ f3 march_mandelbulber(f3 dr, f3 rp, int ni, const FptRenderConfig &cfg, Work &w, int &steps,
                                 bool *found_out = nullptr) {
    const f3 start = rp;
    float d = 0.0f, thr = mb_threshold(rp, cfg), step = 0.0f;
    bool found = false;
    const int max_iter = ni < 10000 ? ni : 10000;
    steps = 0;
    for (int i = 0; i < max_iter; ++i) {
        thr = mb_threshold(rp, cfg);
        d = map_kifs(rp, cfg, w);
        ++steps;
        if (!isfinite(d)) break;
        if (d < thr) {
            found = true;
            break;
        }
        step = mb_step(d, thr, cfg);
        const f3 nx = rp + dr * step;
        if (same3(nx, rp)) break;
        rp = nx;
        if (length(rp - start) >= cfg.render[4]) break;
    }
    if (found_out) *found_out = found;
    if (!found) return rp;
    const float limit = 1.0f - 0.001f * fmaxf(cfg.vset_values[129], 0.0f);
    step *= 0.5f;
    for (int r = 0; r < 30; ++r) {
        if (d < thr && d > thr * limit) break;
        if (d > thr) {
            const f3 nx = rp + dr * step;
            if (same3(nx, rp)) break;
            rp = nx;
        } else if (d < thr * limit) {
            const f3 nx = rp - dr * step;
            if (same3(nx, rp)) break;
            rp = nx;
        }
        d = map_kifs(rp, cfg, w);
        step *= 0.5f;
    }
    return rp;
}

// marchCounted / marchStepCount for Mandelbulber scenes (SM:6857-6908), the
// march of the path-cost diagnostics: the generic loop with the Mandelbulber
// threshold, step and convergence test (d below the LOD-blended threshold),
// no bracketing refinement, at most 10000 steps.
__device__ f3 march_counted_mandelbulber(f3 dr, f3 rp, int ni, float lod_falloff, const FptRenderConfig &cfg,
                                         Work &w, int &steps) {
    const f3 cam = rp;
    const int max_iter = ni < 10000 ? ni : 10000;
    steps = 0;
    for (int i = 0; i < max_iter; ++i) {
        ++steps;
        const float d = map_kifs(rp, cfg, w);
        if (!isfinite(d)) break;
        const float thr = mb_threshold(rp, cfg);
        const float o = mb_step(d, thr, cfg);
        rp = rp + dr * o;
        float lod = thr;
        if (lod_falloff > 0.00001f) {
            const f3 df = cam - rp;
            const float l2 = mixf(lod, 0.1f, dot(df, df) * lod_falloff);
            lod = mixf(thr, l2, cfg.render[5]);
        }
        if (d < lod) break;
        if (cfg.render[4] < o) break;
    }
    return rp;
}

// Normal-stencil field sample: 5x the iterations in iteration-limit mode,
// the raw field distance (no interior rule, no half-step mix).
__device__ __forceinline__ float4 mb_normal_sample(f3 p, const FptRenderConfig &cfg, Work &w) {  // SM:964-974
    w.distance_evals++;
    const float4 s = kifs_field_sample(p, cfg, cfg.vset_values[115] > 1.5f ? 5 : 1);
    w.orbit_iterations += int(fabsf(s.w));
    return s;
}

__device__ f3 normal_mandelbulber(f3 p, const FptRenderConfig &cfg, Work &w) {  // SM:3871-3960
    const float *v = cfg.vset_values;
    if (cfg.render[7] > 0.0f) {
        // slow shading: the iteration-count field on an 11^3 stencil
        const float delta = length(ld3(cfg.camera_position) - p) * cfg.render[7] * v[114];
        const float max_it = cfg.set_values[1];
        f3 g = mk(0.f, 0.f, 0.f);
        for (int ix = -5; ix <= 5; ++ix)
            for (int iy = -5; iy <= 5; ++iy)
                for (int iz = -5; iz <= 5; ++iz) {
                    const f3 sd = mk(float(ix), float(iy), float(iz)) * 0.2f;
                    w.distance_evals++;
                    const float4 s = kifs_field_sample(p + sd * delta, cfg, 1);
                    w.orbit_iterations += int(fabsf(s.w));
                    g = g + sd * (1.0f + max_it - fabsf(s.w));
                }
        const float m = fmaxf(fabsf(g.x), fmaxf(fabsf(g.y), fabsf(g.z)));
        if (!(m > 0.0f) || !isfinite(m)) return mk(1.f, 0.f, 0.f);
        return normalize(g / m);
    }
    const float e = mb_threshold(p, cfg) * v[114];
    if (cfg.sdf_normal_mode == 1u) {
        const float te = e * 0.57735026919f;
        const f3 k0 = mk(1.f, -1.f, -1.f), k1 = mk(-1.f, -1.f, 1.f), k2 = mk(-1.f, 1.f, -1.f), k3 = mk(1.f, 1.f, 1.f);
        const float d0 = mb_normal_sample(p + k0 * te, cfg, w).x, d1 = mb_normal_sample(p + k1 * te, cfg, w).x;
        const float d2 = mb_normal_sample(p + k2 * te, cfg, w).x, d3 = mb_normal_sample(p + k3 * te, cfg, w).x;
        return normalize_fd(k0 * d0 + k1 * d1 + k2 * d2 + k3 * d3);
    }
    const f3 ex = mk(e, 0.f, 0.f), ey = mk(0.f, e, 0.f), ez = mk(0.f, 0.f, e);
    f3 n = mk(mb_normal_sample(p + ex, cfg, w).x - mb_normal_sample(p - ex, cfg, w).x,
              mb_normal_sample(p + ey, cfg, w).x - mb_normal_sample(p - ey, cfg, w).x,
              mb_normal_sample(p + ez, cfg, w).x - mb_normal_sample(p - ez, cfg, w).x);
    if (!(fmaxf(fabsf(n.x), fmaxf(fabsf(n.y), fabsf(n.z))) > 0.0f)) {
        // samples collapsed to one float: the log orbit radius has the same
        // level-set orientation
        auto lr = [&](f3 q) { return logf(fmaxf(fabsf(mb_normal_sample(q, cfg, w).y), 1.0e-30f)); };
        n = mk(lr(p + ex) - lr(p - ex), lr(p + ey) - lr(p - ey), lr(p + ez) - lr(p - ez));
    }
    return normalize_fd(n);
}

__device__ __forceinline__ bool is_mandelbulber(const FptRenderConfig &cfg) {
    return cfg.sdf_id == FPT_SDF_MANDELBULBER;
}

// Scene kinds a kernel is built for: SCENE_ANY decides per config at run
// time; SCENE_SDF (typed programs and built-in fractals without lowered
// tables), SCENE_LOWERED (the same, tables allowed) and SCENE_MB (imported
// Mandelbulber scenes) leave the other kinds' code out, so each kernel
// instance carries only the registers its scenes need.
enum : int { SCENE_ANY = 0, SCENE_SDF = 1, SCENE_MB = 2, SCENE_LOWERED = 3 };
template <int K>
__device__ __forceinline__ bool mb_scene(const FptRenderConfig &cfg) {
    return K == SCENE_MB || (K == SCENE_ANY && is_mandelbulber(cfg));
}
template <int K>
constexpr bool lowered_kind = K == SCENE_ANY || K == SCENE_LOWERED;

// Entry points for every SDF scene: the Mandelbulber branches above for
// imported scenes, the functions of the SDF path otherwise.
template <int K = SCENE_ANY>
__device__ __forceinline__ f3 march_any(f3 dr, f3 rp, int ni, float min_dist, float lod_falloff,
                                        const FptRenderConfig &cfg, Work &w, int &steps) {  // SM:3753-3760
    if (mb_scene<K>(cfg)) return march_mandelbulber(dr, rp, ni, cfg, w, steps);
    return march<lowered_kind<K>>(dr, rp, ni, min_dist, lod_falloff, cfg, w, steps);
}
// The path-cost diagnostics' march (SM:6857-6908): the scene's march,
// except that Mandelbulber scenes use the generic loop with their step rules.
template <int K = SCENE_ANY>
__device__ __forceinline__ f3 march_counted_any(f3 dr, f3 rp, int ni, float min_dist, float lod_falloff,
                                                const FptRenderConfig &cfg, Work &w, int &steps) {
    if (mb_scene<K>(cfg)) return march_counted_mandelbulber(dr, rp, ni, lod_falloff, cfg, w, steps);
    return march<lowered_kind<K>>(dr, rp, ni, min_dist, lod_falloff, cfg, w, steps);
}
template <int K = SCENE_ANY>
__device__ __forceinline__ f3 normal_any(f3 p, const FptRenderConfig &cfg, Work &w) {
    if (mb_scene<K>(cfg)) return normal_mandelbulber(p, cfg, w);
    return lowered_kind<K> ? normal_scene(p, cfg, w) : normal_at(p, cfg, w);
}
template <int K = SCENE_ANY>
__device__ __forceinline__ float map_any(f3 p, const FptRenderConfig &cfg, Work &w) {
    if (mb_scene<K>(cfg)) return map_kifs(p, cfg, w);
    return lowered_kind<K> ? map_scene(p, cfg, w) : map_sdf(p, cfg, w);
}
template <int K = SCENE_ANY>
__device__ __forceinline__ Material material_any(f3 p, const FptRenderConfig &cfg, Work &w) {  // SM:3539-3562
    if (!mb_scene<K>(cfg)) return user_material(p, cfg, w);
    float orbit = 0.0f;
    // a solid style does not read the orbit
    if (cfg.fractal_style_mode != 2u) de_kifs(p, cfg, &orbit, w);
    return material_at_param(material_param(orbit, cfg), cfg);
}

// estimateFocusDistance (SM:5869-5878) for every SDF scene: the centre ray
// of the pinhole camera, marched as the scene marches.
__device__ float focus_estimate_any(const FptRenderConfig &cfg, Work &w) {
    if (!is_mandelbulber(cfg)) return focus_estimate(cfg, w);
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    const f3 dr = rotate_camera_roll(normalize(mk(0.f, 0.f, focal)), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1],
                                     cfg.camera_roll);
    const f3 cam = ld3(cfg.camera_position);
    int steps = 0;
    const f3 hit = march_mandelbulber(dr, cam, 120, cfg, w, steps);
    const float d = length(hit - cam);
    if (!isfinite(d) || d <= 0.001f || d > cfg.render[4] * 0.99f) return 5.0f;
    return d;
}

// The camera ray through screen point (x, y) (SM:5938-5949); false when a
// cut fish-eye does not see it.
template <int K = SCENE_ANY>
__device__ __forceinline__ bool camera_ray_any(float x, float y, const FptRenderConfig &cfg, f3 &dr) {
    if (mb_scene<K>(cfg)) {
        if (!mb_projection_visible(x, y, cfg)) return false;
        dr = mb_camera_ray(x, y, cfg);
        return true;
    }
    const float focal = 1.0f / tanf(cfg.camera_fov / 2.0f * kPi / 180.0f);
    dr = rotate_camera_roll(normalize(mk(x, y, focal)), cfg.camera_yaw_pitch[0], cfg.camera_yaw_pitch[1],
                            cfg.camera_roll);
    return true;
}

// sdfScreenUv (SM:6847-6855): Mandelbulber scenes sample pixel corners
// offset (0, 1), the others pixel centres.
template <int K = SCENE_ANY>
__device__ __forceinline__ UV screen_uv_any(uint32_t x, uint32_t y, const FptRenderConfig &cfg) {
    if (!mb_scene<K>(cfg)) return screen_uv(x, y, cfg.width, cfg.height);
    const float fw = float(cfg.width), fh = float(cfg.height);
    const float u = __fsub_rn(__fdiv_rn(float(x), fw), 0.5f);
    const float v = __fsub_rn(__fdiv_rn(__fadd_rn(float(y), 1.0f), fh), 0.5f);
    return {__fmul_rn(u, __fdiv_rn(fw, fh)), v};
}

}  // namespace fpt
