"""Voxel leaf refinement (Shaders.metal:4312-4490) in the oracle.

Pins:
  - the reference's known answers (contract bits 22-26, Shaders.metal:
    5309-5341) in FP64 and FP32: on the unit sphere the x axis from
    (-2, 0, 0) is refined to t = 1 within 1e-4 by secant/bisection and by
    the restricted trace, the line y = 1.5 has no crossing (both decline),
    and the fixed-DE method declines a typed program;
  - analytic intersections: random rays through a sphere (also scaled and
    translated by the program) refine to the closed-form entry t;
  - fixed DE on the built-in Mandelbox: an accepted t has |DE| no larger
    than every one of its 13 scan samples (brute force) and within the
    acceptance bound max(eps, 1e-5, span / 64).
"""
import math

import numpy as np
import pytest

from fractalpathtracer_amd import ffi, scenes
from oracle import bind


def _sphere_program(radius=1.0, pre=()):
    c = scenes.default_config()
    scenes.set_program(c, list(pre) + [(ffi.OP_SPHERE, 0, [radius])])
    c.voxel_surface_band = 0.5
    return c


@pytest.mark.parametrize("prec", [64, 32])
def test_contract_known_answers(prec):
    c = _sphere_program()
    x, y, dx = (-2.0, 0.0, 0.0), (-2.0, 1.5, 0.0), (1.0, 0.0, 0.0)
    ok, t = bind.leaf_refine(c, 1, x, dx, 0.75, 1.25, precision=prec)
    assert ok and 0.75 <= t <= 1.25 and abs(t - 1.0) <= 1e-4  # bit 22
    assert not bind.leaf_refine(c, 1, y, dx, 0.75, 1.25, precision=prec)[0]  # bit 23
    ok, t = bind.leaf_refine(c, 2, x, dx, 0.75, 1.25, (-1, 0, 0), (0.25,) * 3, precision=prec)
    assert ok and 0.75 <= t <= 1.25 and abs(t - 1.0) <= 1e-4  # bit 24
    assert not bind.leaf_refine(c, 2, y, dx, 0.75, 1.25, (-1, 1.5, 0), (0.25,) * 3, precision=prec)[0]  # bit 25
    assert not bind.leaf_refine(c, 3, x, dx, 0.75, 1.25, precision=prec)[0]  # bit 26


def _sphere_hits(n, seed, centre, radius):
    """Rays from outside towards a sphere: (origin, unit dir, entry t)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        o = centre + rng.normal(size=3) * 3.0
        if np.linalg.norm(o - centre) < radius * 1.5:
            continue
        target = centre + rng.normal(size=3) * radius * 0.15  # near-normal incidence
        d = (target - o) / np.linalg.norm(target - o)
        b = d @ (o - centre)
        disc = b * b - ((o - centre) @ (o - centre) - radius * radius)
        if disc <= 1e-6:
            continue
        out.append((o, d, -b - math.sqrt(disc)))
    return out


@pytest.mark.parametrize("pre,centre,radius", [
    ((), np.zeros(3), 1.0),
    (((ffi.OP_TRANSLATE, 0, [0.3, -0.2, 0.1]),), np.array([0.3, -0.2, 0.1]), 0.8),
    (((ffi.OP_SCALE, 0, [2.0]),), np.zeros(3), 0.35),  # p * 2: the sphere of radius 0.7 shrinks to 0.35
])
def test_refinement_finds_the_analytic_sphere_entry(pre, centre, radius):
    c = _sphere_program(0.7 if pre and pre[0][0] == ffi.OP_SCALE else radius, pre)
    for o, d, t_star in _sphere_hits(24, 5, centre, radius):
        cell = 0.05
        t0, t1 = t_star - 0.4 * cell, t_star + 0.6 * cell
        mid = o + d * (0.5 * (t0 + t1))
        ok, t = bind.leaf_refine(c, 1, o, d, t0, t1)
        assert ok and abs(t - t_star) <= 1e-5, (t, t_star)
        ok, t = bind.leaf_refine(c, 2, o, d, t0, t1, mid, (cell,) * 3)
        # steps of 0.99 d shrink the gap by 1 - 0.99 cos(incidence) <= 0.45 here, so 12 steps from
        # 0.02 away reach the tolerance max(1e-6, 1e-4 (t1 - t0)) = 5e-6 in distance
        cos_i = abs(d @ ((o + d * t_star - centre) / radius))
        assert cos_i > 0.55
        assert ok and abs(t - t_star) <= 5e-6 / cos_i * 1.01 + 1e-9


def _march(c, o, d, eps=1e-5, steps=4000):
    """FP64 sphere tracing to the first surface point (None on a miss)."""
    t = 0.0
    for _ in range(steps):
        dist = bind.map_distance(c, (o + t * d)[None, :], 64)[0]
        if abs(dist) < eps:
            return t
        t += 0.99 * abs(dist)
        if t > 60:
            return None
    return None


def test_fixed_de_on_the_builtin_mandelbox():
    c = scenes.voxel_mandelbox_scene(16, 16, 1, resolution=64)
    rng = np.random.default_rng(9)
    accepted = 0
    for _ in range(30):
        o = rng.normal(size=3) * 9.0
        d = -o / np.linalg.norm(o) + rng.normal(size=3) * 0.05
        d /= np.linalg.norm(d)
        tc = _march(c, o, d)
        if tc is None:
            continue
        # a cell interval around the crossing, as the voxel hit of that cell gives
        t0, t1 = tc - 0.03, tc + 0.05
        ok, t = bind.leaf_refine(c, 3, o, d, t0, t1)
        scan = np.abs(bind.map_distance(c, o + (t0 + np.arange(13)[:, None] * (t1 - t0) / 12) * d, 64))
        accept = max(c.render[3], 1e-5, (t1 - t0) / 64)
        if not ok:
            assert scan.min() > accept  # the refinement only ever lowers the best scan value
            continue
        accepted += 1
        de = abs(bind.map_distance(c, (o + t * d)[None, :], 64)[0])
        assert t0 <= t <= t1 and de <= scan.min() + 1e-12 and de <= accept
    assert accepted >= 10
    # a typed program is not a built-in fractal: declined
    assert not bind.leaf_refine(_sphere_program(), 3, (-2, 0, 0), (1, 0, 0), 0.9, 1.1)[0]
