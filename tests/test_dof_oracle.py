"""Depth of field of the SDF camera (Shaders.metal:5950-5956) and the focus
prepass (estimateFocusDistance, Shaders.metal:5869-5878; host prepass
metal_bridge.mm:4827-4838, 5455-5481) in the CPU oracle, pinned by closed
forms:

  - the focus estimate of the sphere fixture is its hit distance (camera at
    z = -3.5, unit sphere: 2.5, less at most the march threshold); a miss, a
    hit beyond 0.99 max_distance and a degenerate hit give 5;
  - thin-lens optics: the lens sample is uniform over a disc of radius
    R = sqrt(2 pi) dof (randomPoint, Shaders.metal:570-576), so a silhouette
    at distance D seen with focus distance F is blurred by a uniform disc of
    screen radius b = f R |1/D - 1/F| (f = 1 / tan(fov / 2)); its 10%-90%
    coverage width is 2 x b, x from the circular-segment area, and focusing
    on the silhouette (F = D) keeps it as sharp as the pinhole camera.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from fractalpathtracer_amd import scenes
from oracle import bind

GOLDEN = Path(__file__).resolve().parent / "golden"


def _sphere():
    return scenes.apply_scene_json(scenes.default_config(), GOLDEN / "scenes" / "Exact_Sphere.json")


def test_focus_estimate_is_the_sphere_hit_distance():
    c = _sphere()
    assert c.camera_position[2] == -3.5
    # march threshold 0.001 with LOD falloff 0.001 and adaptive marching 0.18:
    # lod <= mix(0.001, 0.1, 0.001 * 2.5^2) -> the stop is within 2e-3 of 2.5
    for prec in (64, 32):
        f = bind.focus_estimate(c, prec)
        assert 2.5 - 2e-3 <= f <= 2.5, (prec, f)
    assert abs(bind.focus_estimate(c, 32) - bind.focus_estimate(c, 64)) < 1e-5


def test_focus_estimate_fallbacks():
    c = _sphere()
    c.camera_yaw_pitch[0] = math.pi  # looking away: the march leaves through max_distance
    assert bind.focus_estimate(c, 64) == 5.0
    c = _sphere()
    c.render[4] = 2.0  # first step 0.99 * 2.5 > max_distance: stops at 2.475 > 0.99 * 2
    assert bind.focus_estimate(c, 64) == 5.0
    c = _sphere()
    c.camera_position[2] = -1.0005  # starts within the threshold of the surface: travel <= 0.001
    assert bind.focus_estimate(c, 64) == 5.0


def _coverage_scene(dof, focus, n=96, spp=64):
    """Unit sphere seen from z = -6 as an emissive disc on black: the image is
    the fraction of samples that hit (sun, environment and gradient off, one
    path segment)."""
    c = _sphere()
    c.width = c.height = n
    c.samples = spp
    c.camera_position[2] = -6.0
    c.camera_dof, c.focus_distance = dof, focus
    c.render[0] = 1.0
    c.render[5] = 0.0
    c.sun[0] = 0.0
    c.world[4] = 0.0
    c.world[6] = 0.0
    for i in range(3):
        c.program_material[i] = 1.0
    c.program_material[7] = 1.0
    return c


def _edge(img, n):
    """Radial coverage profile around the image centre: (10%-90% width, 50%
    radius) in pixels, at 0.5 px resolution."""
    cov = img[..., 0] / img[n // 2, n // 2, 0]
    y, x = np.mgrid[0:n, 0:n]
    r = np.hypot(x + 0.5 - n / 2, y + 0.5 - n / 2).ravel()
    edges = np.arange(0, n / 2, 0.5)
    idx = np.digitize(r, edges)  # bin i holds radii in [edges[i-1], edges[i])
    sums = np.bincount(idx, weights=cov.ravel(), minlength=len(edges) + 1)[1:len(edges)]
    counts = np.bincount(idx, minlength=len(edges) + 1)[1:len(edges)]
    prof = np.where(counts > 0, sums / np.maximum(counts, 1), 1.0)  # the empty innermost bin is inside
    rc = edges[:-1] + 0.25
    return rc[np.argmax(prof < 0.1)] - rc[np.argmax(prof < 0.9)], rc[np.argmax(prof < 0.5)]


def _segment_fraction_point(p):
    """x in [0, 1] where a straight edge at offset x b covers fraction p of a
    uniform disc of radius b: (acos x - x sqrt(1 - x^2)) / pi = p."""
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = (lo + hi) / 2
        if (math.acos(mid) - mid * math.sqrt(1 - mid * mid)) / math.pi > p:
            lo = mid
        else:
            hi = mid
    return lo


def test_thin_lens_blur_and_focus_plane():
    n = 96
    f = 1.0 / math.tan(math.radians(52 / 2))
    dist = math.sqrt(6.0 ** 2 - 1.0)  # camera to the silhouette points
    r_sil = f * math.tan(math.asin(1.0 / 6.0)) * n  # silhouette radius in pixels
    x10 = _segment_fraction_point(0.1)
    w_pin, r_pin = _edge(bind.render(_coverage_scene(0.0, 0.0, n), 64), n)
    assert w_pin <= 1.5 and abs(r_pin - r_sil) <= 1.0
    # focused on the silhouette: as sharp as the pinhole whatever the aperture
    w_sharp, r_sharp = _edge(bind.render(_coverage_scene(0.1, dist, n), 64), n)
    assert w_sharp <= w_pin + 0.5 and abs(r_sharp - r_sil) <= 1.0
    widths = []
    for dof in (0.05, 0.1):
        w, r50 = _edge(bind.render(_coverage_scene(dof, 2.0, n), 64), n)
        b = f * math.sqrt(2 * math.pi) * dof * abs(1 / dist - 1 / 2.0) * n
        assert w == pytest.approx(2 * x10 * b, rel=0.15), (dof, w, b)
        assert abs(r50 - r_sil) <= 0.05 * b + 1.0  # blur keeps the edge centred
        widths.append(w)
    assert widths[1] / widths[0] == pytest.approx(2.0, rel=0.15)


def test_dof_in_focus_centre_pixel_keeps_direct_light_closed_form():
    """With the focus at the estimated hit distance every lens ray of the
    centre pixel converges on the same surface point, so its direct light is
    the pinhole closed form (sun * max(n.l, 0) * roughness * albedo)."""
    c = _sphere()
    c.width, c.height, c.samples = 9, 9, 8
    c.render[0] = 1.0
    c.render[5] = 0.0
    c.sun[4] = 0.0
    c.world[4] = 0.0
    c.program_material[7] = 0.0
    c.camera_dof, c.focus_distance = 0.2, 0.0
    img = bind.render(c, 64)
    l_dir = np.array([math.cos(math.radians(c.sun[2])) * math.sin(math.radians(c.sun[1])),
                      math.sin(math.radians(c.sun[2])),
                      math.cos(math.radians(c.sun[2])) * math.cos(math.radians(c.sun[1]))])
    n = np.array([0.0, 0.0, -1.0])
    want = c.sun[3] * max(n @ l_dir, 0.0) * c.program_material[3] * np.array(c.sun_color) * np.array(
        c.program_material[:3])
    assert np.allclose(img[4, 4, :3], want, rtol=5e-3)
    # defocused (focus in front of the sphere): the centre pixel sees other points
    c.focus_distance = 1.0
    off = bind.render(c, 64)
    assert not np.allclose(off[4, 4, :3], want, rtol=5e-3)


def test_zero_aperture_is_the_pinhole_camera():
    c = _sphere()
    c.width, c.height, c.samples = 24, 16, 2
    a = bind.render(c, 64)
    c.camera_dof = 1e-7
    b = bind.render(c, 64)
    assert np.sqrt(np.mean((a - b) ** 2)) < 1e-4
