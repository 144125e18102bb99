"""Pins of the CPU oracle (no GPU).

Each test checks an oracle function against something that does not come
from the oracle's own formula: the reference's golden vectors
(src/mandelbulber/orbit.rs:1866-1960), closed forms, invariants, or FP64
central differences (brute force).
"""
import math

import numpy as np
import pytest

from fractalpathtracer_amd import ffi, scenes
from oracle import bind

GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"


# --- RNG (Shaders.metal:541-576) ------------------------------------------------

def _hash13_numpy(x, y, z):
    """Second, independent restatement of hash13bits (Shaders.metal:541-551)
    in vectorised uint32 numpy arithmetic."""
    m = np.uint32(0x5BD1E995)
    b = [np.asarray(v, np.float32).view(np.uint32) for v in (x, y, z)]
    with np.errstate(over="ignore"):
        h = np.full(b[0].shape, 1190494759, np.uint32)
        for bits in b:
            h = (h ^ bits) * m
        h ^= h >> np.uint32(13)
        h = h * m
        h ^= h >> np.uint32(15)
    return h


def test_hash13_bits_match_independent_restatement():
    rng = np.random.default_rng(1)
    pts = rng.uniform(-1, 1, (500, 3)).astype(np.float32)
    want = _hash13_numpy(pts[:, 0], pts[:, 1], pts[:, 2])
    got = np.array([bind.hash13_bits(*map(float, p)) for p in pts], np.uint32)
    assert np.array_equal(got, want)


def test_hash13_is_24bit_uniform():
    rng = np.random.default_rng(2)
    pts = rng.uniform(-0.8, 0.8, (4000, 2)).astype(np.float32)
    vals = np.array([bind.hash13(float(a), float(b), 3.0) for a, b in pts])
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert np.all(np.round(vals * 2 ** 24) == vals * 2 ** 24)
    hist, _ = np.histogram(vals, bins=16, range=(0, 1))
    chi2 = float(((hist - 250.0) ** 2 / 250.0).sum())
    assert chi2 < 40.0  # 15 dof, p ~ 5e-4


# --- Formula kernels: reference golden vectors -----------------------------------

def test_mandelbulb_power2_step_golden_orbit_rs_1876():
    """classic_mandelbulb_kernel_matches_the_source_equations (orbit.rs:1876-1888)."""
    z, dr = bind.mandelbulb_step(np.array([1.0, 2.0, 3.0]), power=2.0)
    r = math.sqrt(14.0)
    assert abs(dr - (2.0 * r + 1.0)) < 1e-12
    assert abs(np.linalg.norm(z) - 14.0) < 1e-12


def test_mandelbulb_power8_closed_form():
    """|z'| = r^p and the azimuth multiplies by p (triplex power)."""
    p0 = np.array([0.3, -0.4, 0.5])
    z, dr = bind.mandelbulb_step(p0, power=8.0)
    r = np.linalg.norm(p0)
    assert abs(np.linalg.norm(z) - r ** 8) < 1e-14
    assert abs(dr - (8 * r ** 7 + 1)) < 1e-14
    phi = math.atan2(p0[1], p0[0])
    assert abs(math.remainder(math.atan2(z[1], z[0]) - 8 * phi, 2 * math.pi)) < 1e-9
    theta = math.asin(p0[2] / r)
    assert abs(z[2] / np.linalg.norm(z) - math.sin(8 * theta)) < 1e-12


def test_mandelbox_step_golden_orbit_rs_1915():
    """mandelbox_fast_uses_derived_radius_factors / full_mandelbox (orbit.rs:1915-1960)."""
    z, dr = bind.mandelbox_step(np.array([0.1, 0.2, 0.3]), scale=-1.5, min_r2=0.25, fixed_r2=1.0)
    assert np.allclose(z, [-0.6, -1.2, -1.8], atol=1e-12)
    assert dr == 7.0


def test_mandelbox_fixed_radius_and_fold_branches():
    # r^2 = 0.5 in (minR^2, fixedR^2): factor fixedR^2/r^2 = 2
    z, dr = bind.mandelbox_step(np.array([0.5, 0.5, 0.0]), scale=2.0, min_r2=0.25, fixed_r2=1.0)
    assert np.allclose(z, [2.0, 2.0, 0.0]) and dr == 2.0 * 2.0 + 1.0
    # box fold: |x| > 1 -> sign(x)*2 - x, then r^2 = 1.22 >= fixedR^2: no sphere fold
    z, dr = bind.mandelbox_step(np.array([1.5, -1.6, 0.9]), scale=1.0, min_r2=0.25, fixed_r2=1.0)
    assert np.allclose(z, [0.5, -0.4, 0.9]) and dr == 2.0


# --- Leaves ------------------------------------------------------------------------

def test_mandelbulb_leaf_escapes_on_first_iteration_far_away():
    p = np.array([0.0, 0.0, -3.0])
    out = bind.leaf_orbit(ffi.OP_MANDELBULB, [8, 12, 2, 0], p)
    # theta = -pi/2 -> 8 theta = -4 pi: z1 = (3^8, 0, 0) + p
    z1 = np.array([3.0 ** 8, 0.0, -3.0])
    r1 = np.linalg.norm(z1)
    dr1 = 8 * 3.0 ** 7 + 1
    assert out["iterations"] == 1 and out["escaped"]
    assert abs(out["distance"] - 0.5 * r1 * math.log(r1) / dr1) < 1e-12
    assert np.allclose(out["z"], z1, rtol=1e-12)


def test_mandelbulb_leaf_interior_is_zero_distance():
    out = bind.leaf_orbit(ffi.OP_MANDELBULB, [8, 12, 2, 0], [0.1, 0.05, -0.02])
    assert out["iterations"] == 12 and not out["escaped"] and out["radius"] <= 1.0
    assert out["distance"] == 0.0


def test_mandelbox_leaf_linear_finalizer_matches_orbit_steps():
    """Leaf = orbit loop of mandelbox steps + c with bailout 100 and r/dr."""
    p = np.array([0.7, -0.3, 0.45])
    z, dr = p.copy(), 1.0
    for it in range(10):
        z, dr = bind.mandelbox_step(z, -1.5, 0.25, 1.0, dr)
        z = z + p
        if np.linalg.norm(z) > 100:
            break
    out = bind.leaf_orbit(ffi.OP_MANDELBOX, [-1.5, 10, 0.5, 1.0], p)
    assert abs(out["distance"] - np.linalg.norm(z) / dr) < 1e-12 * max(1, abs(out["distance"]))


def test_menger_leaf_known_points():
    # Menger sponge (size 1, divisor 3): the centre is removed by the first cut;
    # a point outside the cube is at its box distance.
    d_centre = bind.leaf_orbit(ffi.OP_MENGER, [5, 1, 3, 0], [0.0, 0.0, 0.0])["distance"]
    assert abs(d_centre - 1.0 / 3.0) < 1e-12   # distance to the cross boundary
    d_out = bind.leaf_orbit(ffi.OP_MENGER, [5, 1, 3, 0], [2.5, 0.9, 0.9])["distance"]
    assert abs(d_out - 1.5) < 1e-12
    # a corner cube of the first level is solid at its centre (2/3, 2/3, 2/3)
    d_corner = bind.leaf_orbit(ffi.OP_MENGER, [1, 1, 3, 0], [2 / 3, 2 / 3, 2 / 3])["distance"]
    assert d_corner < 0.0


# --- Program interpreter: closed forms of the bundled exact fixtures ---------------

def _scene(name):
    c = scenes.default_config()
    return scenes.apply_scene_json(c, GOLDEN / "scenes" / f"{name}.json")


def test_exact_sphere_distance_and_normal():
    c = _scene("Exact_Sphere")
    pts = np.array([[0, 0, -3.5], [2.0, 0, 0], [0.3, 0.2, 0.1]])
    d = bind.map_distance(c, pts)
    assert np.allclose(d, np.linalg.norm(pts, axis=1) - 1.0, atol=1e-15)
    n = bind.normal(c, np.array([[0, 0, -1.0], [0.6, 0.8, 0]]))
    assert np.allclose(n, [[0, 0, -1], [0.6, 0.8, 0]], atol=1e-12)


def _box(p, b):
    q = np.abs(p) - b
    return np.linalg.norm(np.maximum(q, 0)) + min(q.max(), 0.0)


def test_exact_csg_box_minus_sphere():
    c = _scene("Exact_CSG")
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1.5, 1.5, (200, 3))
    # program data are float32 (FptSdfInstruction.data)
    b = np.array([1.1, 0.85, 0.65], np.float32).astype(np.float64)
    r = float(np.float32(0.48))
    want = [max(_box(p, b), -(np.linalg.norm(p) - r)) for p in pts]
    assert np.allclose(bind.map_distance(c, pts), want, atol=1e-12)


def test_program_gradient_matches_fp64_central_differences():
    rng = np.random.default_rng(4)
    for cfg in (scenes.mandelbulb_scene(8, 8), scenes.typed_program_scene(8, 8)):
        pts = rng.uniform(-1.3, 1.3, (60, 3))
        surf = bind.program_surface(cfg, pts)
        h = 1e-7
        for p, s in zip(pts, surf):
            g = np.array([(bind.map_distance(cfg, [p + h * e])[0] - bind.map_distance(cfg, [p - h * e])[0]) / (2 * h)
                          for e in np.eye(3)])
            if np.linalg.norm(g) < 1e-6:
                continue
            # skip points within h of a fold / CSG seam (gradient is discontinuous there)
            if abs(s[0] - bind.map_distance(cfg, [p])[0]) > 1e-12:
                continue
            err = np.linalg.norm(s[1:] - g) / np.linalg.norm(g)
            assert err < 1e-4 or _near_seam(cfg, p), (p, s, g)


def _near_seam(cfg, p):
    """True when the FP64 gradient is not smooth at p (two one-sided
    differences disagree): fold planes, CSG seams, box-fold boundaries."""
    g = []
    for h in (1e-6, 1e-5):
        g.append([(bind.map_distance(cfg, [p + h * e])[0] - bind.map_distance(cfg, [p])[0]) / h for e in np.eye(3)])
    g = np.array(g)
    return np.linalg.norm(g[0] - g[1]) > 1e-3 * max(np.linalg.norm(g[1]), 1e-9)


# --- March and render ---------------------------------------------------------------

def test_primary_march_hits_sphere_at_closed_form_t():
    c = _scene("Exact_Sphere")
    c.width, c.height = 33, 17
    c.render[5] = 0.0
    tr = bind.trace_primary(c, 64)
    centre = tr[8, 16]
    assert centre[7] == 1.0
    eps = c.render[3]
    # advance-then-threshold: stops once 0.99 d < eps, so t is within eps of 2.5
    assert 2.5 - eps <= centre[0] <= 2.5 + 1e-12
    assert np.allclose(centre[4:7], [0, 0, -1], atol=1e-3)
    # invariant: every hit ends within min_distance of the surface
    hits = tr[..., 7] == 1
    dist_at_end = np.abs(np.linalg.norm(tr[..., 1:4][hits], axis=-1) - 1.0)
    assert dist_at_end.max() < eps


def test_direct_light_closed_form_on_sphere():
    """render[0] = 1: col = sun * max(n.l, 0) * roughness * albedo at the centre."""
    c = _scene("Exact_Sphere")
    c.width, c.height, c.samples = 9, 9, 1
    c.render[0] = 1.0
    c.render[5] = 0.0
    c.sun[4] = 0.0          # hard sun
    c.world[4] = 0.0        # no environment
    c.program_material[7] = 0.0  # no emission
    img = bind.render(c, 64)
    l_dir = np.array([math.cos(math.radians(c.sun[2])) * math.sin(math.radians(c.sun[1])),
                      math.sin(math.radians(c.sun[2])),
                      math.cos(math.radians(c.sun[2])) * math.cos(math.radians(c.sun[1]))])
    n = np.array([0.0, 0.0, -1.0])
    want = c.sun[3] * max(n @ l_dir, 0.0) * c.program_material[3] * np.array(c.sun_color) * np.array(
        c.program_material[:3])
    assert np.allclose(img[4, 4, :3], want, rtol=2e-3)


def test_fp32_twin_tracks_fp64_on_smooth_scene():
    c = _scene("Exact_CSG")
    c.width, c.height, c.samples = 48, 27, 2
    a = bind.render(c, 64)
    b = bind.render(c, 32)
    assert np.sqrt(np.mean((a - b) ** 2)) < 1e-3


def test_counters_are_consistent():
    c = scenes.mandelbulb_scene(16, 16, 1)
    bind.render(c, 64)
    k = bind.counters()
    assert k["paths"] == 256
    assert k["distance_evals"] >= k["primary_steps"] + k["secondary_steps"] + k["shadow_steps"]
    assert k["orbit_iterations"] >= k["distance_evals"]


def test_batch_scene_generator_is_seeded_and_in_range():
    a = scenes.batch_scenes(count=64, seed=0)
    b = scenes.batch_scenes(count=64, seed=0)
    c = scenes.batch_scenes(count=64, seed=1)
    assert bytes(a[5]) == bytes(b[5]) and bytes(a[5]) != bytes(c[5])
    for i, s in enumerate(a):
        assert 64 <= s.width <= 512 and 64 <= s.height <= 512 and 1 <= s.samples <= 16
        op = s.sdf_program[s.sdf_program_count - 1]
        if i % 3 == 0:
            assert op.opcode == ffi.OP_MANDELBULB and 2.0 <= op.data[0] <= 10.0
        elif i % 3 == 1:
            assert op.opcode == ffi.OP_MANDELBOX and -2.5 <= op.data[0] <= 2.5
        else:
            assert op.opcode == ffi.OP_MENGER and 3 <= op.data[0] <= 6


def test_sample_range_render_composes():
    """orc_render_samples over [0, spp) is orc_render; the running means of
    [0, k) and [k, n) combine to the mean of [0, n); rows outside the band
    stay zero."""
    c = scenes.mandelbulb_scene(12, 10, 6, bounces=1)
    full = bind.render(c, 64)
    assert np.array_equal(bind.render_samples(c, (0, 6)), full)
    a = bind.render_samples(c, (0, 2))
    b = bind.render_samples(c, (2, 6))
    assert np.allclose((2 * a + 4 * b) / 6, full, atol=1e-12)
    band = bind.render_samples(c, (3, 5), rows=(4, 6))
    assert np.all(band[:4] == 0) and np.all(band[6:] == 0)
    assert np.array_equal(band[4:6], bind.render_samples(c, (3, 5))[4:6])
