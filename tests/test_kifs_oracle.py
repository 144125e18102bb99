"""Pins of the oracle's Mandelbulber native KIFS field (formula 10,
Shaders.metal:786-949, src/mandelbulber/orbit.rs:699-750, Ifs finalizer
orbit.rs:662) by closed forms: the unfolded orbit z_n = k^n (z_0 - o) + o,
abs folds and plane reflections after one iteration, quarter-turn rotation
orders, the global repeat and the screen iteration budget."""
import math

import numpy as np
import pytest

from fractalpathtracer_amd import ffi, scenes
from oracle import bind


def kifs(**kw):
    c = scenes.default_config()
    c.sdf_id = ffi.SDF_MANDELBULBER
    sv = [0.0] * 40
    sv[1], sv[2], sv[3] = kw.get("iterations", 1), kw.get("bailout", 1e6), kw.get("scale", 2.0)
    sv[4], sv[5] = kw.get("abs_mask", 0), kw.get("plane_mask", 0)
    sv[6:9] = kw.get("offset", [0, 0, 0])
    sv[9:12] = kw.get("rotation", [0, 0, 0])
    for k, d in enumerate(kw.get("planes", [])):
        sv[12 + 3 * k:15 + 3 * k] = d
    sv[0] = kw.get("world_scale", 0.0)
    for i, v in enumerate(sv):
        c.set_values[i] = v
    for i, v in kw.get("vset", {}).items():
        c.vset_values[i] = v
    return c


def swap(p):  # the field works on (x, z, y)
    return np.array([p[0], p[2], p[1]], np.float64)


def test_unfolded_orbit_closed_form():
    k, o, n = 2.0, np.array([0.5, 0.25, 0.0]), 5
    c = kifs(iterations=n, scale=k, offset=o)
    p = np.array([0.3, -0.2, 0.7])
    out = bind.kifs_sample(c, [p])[0]
    zn = k ** n * (swap(p) - o) + o
    assert out[1] == pytest.approx(np.linalg.norm(zn), rel=1e-12)
    assert out[2] == k ** n and out[3] == n            # not escaped
    assert out[0] == pytest.approx((np.linalg.norm(zn) - 2) / k ** n, rel=1e-12)


def test_escape_iteration_and_world_scale():
    k, o = 2.0, np.zeros(3)
    c = kifs(iterations=30, scale=k, bailout=3.0, world_scale=2.0)
    p = np.array([0.4, 0.1, -0.2])
    out = bind.kifs_sample(c, [p])[0]
    z0 = swap(p) / 2.0
    n = next(i for i in range(1, 30) if np.linalg.norm(k ** i * z0) > 3.0)
    assert out[3] == -n and out[2] == k ** n
    assert out[0] == pytest.approx((np.linalg.norm(k ** n * z0) - 2) / k ** n * 2.0, rel=1e-12)


def test_abs_fold_and_plane_reflection():
    p = np.array([-0.3, -0.1, 0.2])          # z0 = (-0.3, 0.2, -0.1)
    c = kifs(abs_mask=5)                     # fold x and z
    z = bind.kifs_sample(c, [p])[0]
    assert z[1] == pytest.approx(np.linalg.norm(2 * np.array([0.3, 0.2, 0.1])), rel=1e-12)
    d = (np.array([1.0, 1.0, 0.0]) / math.sqrt(2)).astype(np.float32).astype(np.float64)  # stored as float
    c = kifs(plane_mask=1, planes=[d])       # reflect through the plane x + y = 0
    z0 = swap(p)
    proj = z0 @ d
    refl = z0 - 2 * proj * d if proj < 0 else z0
    assert bind.kifs_sample(c, [p])[0][1] == pytest.approx(np.linalg.norm(2 * refl), rel=1e-12)
    c = kifs(plane_mask=2, planes=[d, -d])   # only plane 1 enabled: z0 . (-d) > 0 -> no reflection
    if z0 @ (-d) >= 0:
        assert bind.kifs_sample(c, [p])[0][1] == pytest.approx(np.linalg.norm(2 * z0), rel=1e-12)


def test_rotation_orders_by_quarter_turns():
    p = np.array([0.3, -0.2, 0.7])
    z0 = swap(p)
    # rotation.x = pi/2 is the last factor Rz: (x, y, z) -> (-y, x, z).
    # Radii are invariant, so test through an offset that breaks symmetry:
    # z1 = 2 * (R (z0 - o) + o) - o = 2 R (z0 - o) + o.
    o = np.array([0.1, 0.2, 0.3])
    for rot, R in [([math.pi / 2, 0, 0], lambda v: np.array([-v[1], v[0], v[2]])),
                   ([0, 0, math.pi / 2], lambda v: np.array([v[0], -v[2], v[1]])),
                   ([0, math.pi / 2, 0], lambda v: np.array([v[2], v[1], -v[0]]))]:
        c = kifs(rotation=rot, offset=o)
        z1 = 2 * R(z0 - o) + o
        assert bind.kifs_sample(c, [p])[0][1] == pytest.approx(np.linalg.norm(z1), rel=1e-6), rot  # pi/2, o in float32


def test_global_repeat_and_rotation():
    # repeat period 2 on the (swapped) x axis: 1.7 -> -0.3
    c = kifs(vset={126: 2.0})
    a = bind.kifs_sample(c, [[1.7, 0.2, 0.1]])[0]
    b = bind.kifs_sample(kifs(), [[-0.3, 0.2, 0.1]])[0]
    assert a[1] == pytest.approx(b[1], rel=1e-12)
    # global rotation2 r.x = pi/2 acts first as Rx on (x, z, y): (x, y', z') -> (x, -z', y')
    c = kifs(vset={123: math.pi / 2})
    p = np.array([0.3, -0.2, 0.7])
    q = swap(p)
    q = np.array([q[0], -q[2], q[1]])
    ref = bind.kifs_sample(kifs(), [swap(q)])[0]
    assert bind.kifs_sample(c, [p])[0][1] == pytest.approx(ref[1], rel=1e-6)


@pytest.mark.parametrize("rate,want", [(0.0, 20), (2.0, 16), (10.0, 15)])
def test_screen_iteration_budget(rate, want):
    """20 configured iterations, footprint 4 over a minimum detail of 1:
    2 octaves, rate r removes floor(2 r) iterations, never below ceil(0.75 * 20)."""
    c = kifs(iterations=20, scale=1.01, vset={130: rate, 115: 0.0, 117: 4.0, 118: 1.0, 119: 100.0})
    assert bind.kifs_sample(c, [[0.01, 0.02, 0.03]])[0][3] == want


def test_fp32_twin_tracks_fp64():
    c = kifs(iterations=12, scale=3.0, bailout=10.0, abs_mask=7, plane_mask=7, offset=[1, 1, 1],
             planes=[[0.7071068, -0.7071068, 0], [0.7071068, 0, -0.7071068], [0, 0.7071068, -0.7071068]])
    pts = np.random.default_rng(0).uniform(-1.5, 1.5, (500, 3))
    a, b = bind.kifs_sample(c, pts, 64), bind.kifs_sample(c, pts, 32)
    same = a[:, 3] == b[:, 3]
    assert same.mean() > 0.97
    assert np.all(np.abs(a[same, 0] - b[same, 0]) <= 5e-5 * (1 + np.abs(a[same, 0])))
