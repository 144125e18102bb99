"""The analytic sky of environment() (world mode 1, Shaders.metal:5764-5787)
in the oracle, pinned by what the model fixes rather than by restating it:

  - it depends on the view only through its elevation and its angle to the
    sun, so turning view and sun together about the vertical leaves it
    unchanged;
  - with the sun at the nadir the sky term is scaled by zero: black
    everywhere;
  - looking at the zenith with the sun there: the Perez luminance
    distribution with Preetham et al.'s turbidity-2 luminance coefficients
    (A = 0.1787 T - 1.4630, B = -0.3554 T + 0.4275, C = -0.0227 T + 5.3251,
    D = 0.1206 T - 2.5771, E = -0.0670 T + 0.3703; "A Practical Analytic
    Model for Daylight", 1999) at theta = gamma = 0, plus the full sun disk;
  - circumsolar brightening: at fixed elevation the sky darkens away from
    the sun (outside the disk);
  - the sun disk has angular radius 0.05 rad: a view 0.04 rad from the sun
    adds 100 (1 - 0.04^2 / 0.05^2 approx.) h^10 over a view just outside.
"""
import math

import numpy as np
import pytest

from fractalpathtracer_amd import scenes
from oracle import bind


def _sky_config(rot_deg, elev_deg):
    c = scenes.default_config()
    c.world[0], c.world[4], c.world[5] = 1.0, 1.0, 1.0  # sky, power 1, contrast 1
    c.world[2], c.world[3] = rot_deg, elev_deg
    return c


def _sun(rot_deg, elev_deg):
    """rotateCamera((0, 0, 1), yaw, pitch)."""
    y, p = math.radians(rot_deg), math.radians(elev_deg)
    v = np.array([0.0, math.sin(p), math.cos(p)])
    return np.array([v[0] * math.cos(y) + v[2] * math.sin(y), v[1], -v[0] * math.sin(y) + v[2] * math.cos(y)])


def _yaw(v, a):
    c, s = math.cos(a), math.sin(a)
    return np.stack([v[:, 0] * c + v[:, 2] * s, v[:, 1], -v[:, 0] * s + v[:, 2] * c], axis=1)


def _dirs(n, seed):
    d = np.random.default_rng(seed).normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def test_turning_view_and_sun_together_leaves_the_sky_unchanged():
    d = _dirs(256, 0)
    a = bind.environment(_sky_config(30.0, 25.0), d)
    for turn in (17.0, 90.0, 200.0):
        b = bind.environment(_sky_config(30.0 + turn, 25.0), _yaw(d, math.radians(turn)))
        assert np.allclose(a, b, rtol=1e-9, atol=1e-9), turn


def test_sun_at_nadir_gives_a_black_sky():
    assert np.all(bind.environment(_sky_config(0.0, -90.0), _dirs(256, 1)) == 0.0)


def test_zenith_with_the_sun_at_zenith_closed_form():
    T = 2.0
    A, B, C = 0.1787 * T - 1.4630, -0.3554 * T + 0.4275, -0.0227 * T + 5.3251
    D, E = 0.1206 * T - 2.5771, -0.0670 * T + 0.3703
    perez = (1 + A * math.exp(B)) * (1 + C + E)  # theta = gamma = 0
    lo, hi = np.array([0.004, 0.048, 0.253]), np.array([0.8, 0.9, 1.0])
    col = (lo + (hi - lo) * perez) * 1.2  # height of view and sun both 1
    col = (col - 0.5) * 1.2 + 0.5
    want = np.maximum(col * 0.8 + 100.0, 0.0)  # full disk, h^10 = 1
    got = bind.environment(_sky_config(0.0, 90.0), [[0.0, 1.0, 0.0]])[0]
    assert np.allclose(got, want, rtol=1e-6)


def _offset(s, g):
    """The direction at angle g from s, towards the zenith side (rotation
    about s x up)."""
    axis = np.cross(s, [0.0, 1.0, 0.0])
    axis /= np.linalg.norm(axis)
    return s * math.cos(g) + np.cross(axis, s) * math.sin(g)


def test_circumsolar_brightening_and_the_sun_disk():
    rot, elev = 40.0, 30.0
    c = _sky_config(rot, elev)
    s = _sun(rot, elev)

    def at(g, side=-1.0):
        return bind.environment(c, [_offset(s, side * g)])[0]

    # below the sun (towards the horizon, h kept > 0): darker with the angle to the sun
    away = [at(g)[2] for g in (0.2, 0.35, 0.5)]
    assert all(x > y for x, y in zip(away, away[1:]))
    # disk of radius 0.05 rad: 100 h^10 max((cos g - 1) / (1 - cos 0.05) + 1, 0) with h the view's
    # height; the smooth sky term is removed by linear extrapolation from two views outside it
    g = 0.03
    d = _offset(s, -g)
    sky_lin = 2 * at(0.06) - at(0.09)  # sky(0.03) from sky(0.06), sky(0.09)
    disk = 100 * ((d[1] + 1) / 2) ** 10 * ((math.cos(g) - 1) / (1 - math.cos(0.05)) + 1)
    assert np.allclose(at(g) - sky_lin, disk, rtol=0.01)
    assert np.all(at(0.0502) - (2 * at(0.06) - at(0.07)) < 0.01 * disk)  # nothing just outside
