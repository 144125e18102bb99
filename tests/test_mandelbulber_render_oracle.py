"""The oracle's Mandelbulber render branches (Shaders.metal:473-511 camera,
975-995 deMandelbulber, 3627-3751 march, 3871-3960 normal, 6847-6855 pixel
offset), pinned on the CPU.

Pins:
- a closed-form field: a Kaleidoscopic IFS with no folds, offset or
  rotation maps z -> k z, so with N iterations and bailout B > 2 its field
  inside the non-escaping ball |z| <= B / k^N is exactly |p| - R with
  R = 2 ws / k^N, and a lower bound outside.  On it the march's final point
  must sit in the refinement band (threshold * (1 - 0.001 detail),
  threshold) above the sphere, in both threshold modes and with the step
  clamps on; the normals must be radial; the iteration-limit mode must
  report interior points at distance 0 and escaped points below the
  threshold at 1.01 threshold;
- the projections' defining properties: perspective = pinhole with focal
  length 1 / tan(fov / 2); fish-eye: the angle to the view axis is r * fov;
  equirectangular: longitude and latitude are linear in the screen point;
  the cut fish-eye sees nothing beyond 90 degrees off axis;
- the pixel offset (0, 1) of Mandelbulber scenes, at exactly representable
  screen points.
The log-orbit-radius fallback of the central normal stencil (six equal
samples) has no such check: parity unpinned.
"""
import math

import numpy as np
import pytest

from fractalpathtracer_amd import ffi
from oracle import bind
from kifs_sphere import LIMIT, R, sphere_cfg


@pytest.mark.parametrize("prec", [64, 32])
def test_closed_form_field_is_the_sphere_distance(prec):
    c = sphere_cfg()
    rng = np.random.default_rng(1)
    dirs = rng.standard_normal((200, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    radii = rng.uniform(10.0, 6000.0, 200)  # inside the non-escaping ball |p| <= 100 ws / 16 = 6400
    d = bind.map_distance(c, dirs * radii[:, None], prec)
    assert np.allclose(d, radii - R, rtol=0, atol=1e-9 if prec == 64 else 0.05)
    far = dirs * 20000.0  # escapes earlier: a lower bound
    assert np.all(bind.map_distance(c, far, prec) <= 20000.0 - R + 1e-6)


@pytest.mark.parametrize("prec", [64, 32])
def test_march_ends_in_the_refinement_band(prec):
    """Fixed threshold 2: every ray that meets the sphere ends with its
    height above it in (2 * LIMIT, 2), on its own ray."""
    c = sphere_cfg(mode=0, thr=2.0)
    rng = np.random.default_rng(2)
    targets = rng.standard_normal((64, 3))
    targets = targets / np.linalg.norm(targets, axis=1, keepdims=True) * rng.uniform(0, 100, (64, 1))
    cam = np.array([0.0, 0.0, -1000.0])
    dirs = targets - cam
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    out = bind.march(c, np.tile(cam, (64, 1)), dirs, prec)
    p = out[:, :3]
    t = np.einsum("ij,ij->i", p - cam, dirs)
    assert np.allclose(p, cam + dirs * t[:, None], atol=1e-6 if prec == 64 else 2e-3)
    h = np.linalg.norm(p, axis=1) - R
    slack = 1e-9 if prec == 64 else 1e-4
    assert np.all(h > 2.0 * LIMIT - slack) and np.all(h < 2.0 + slack), (h.min(), h.max())


@pytest.mark.parametrize("prec", [64, 32])
def test_distance_scaled_threshold_band(prec):
    """Threshold 0.002 x camera distance: from the camera (threshold 0) the
    first step is the whole distance to the sphere, 872; there d = 0 is below
    the threshold 872 * 0.002, which the refinement then brackets."""
    c = sphere_cfg(mode=1, factor=0.002)
    out = bind.march(c, [[0.0, 0.0, -1000.0]], [[0.0, 0.0, 1.0]], prec)
    thr = 872.0 * float(np.float32(0.002))
    h = -out[0, 2] - R
    slack = 1e-9 if prec == 64 else 1e-4
    assert thr * LIMIT - slack < h < thr + slack
    assert out[0, 3] == 2  # the march steps: the camera point and the sphere point


def test_step_clamps_bound_every_step():
    c = sphere_cfg()
    c.vset_values[108], c.vset_values[109], c.vset_values[110] = 1.0, 0.5, 100.0
    out = bind.march(c, [[0.0, 0.0, -1000.0]], [[0.0, 0.0, 1.0]], 64)
    h = np.linalg.norm(out[0, :3]) - R
    assert 2.0 * 0.998 < h < 2.0
    assert out[0, 3] >= math.ceil((1000 - R - 2.0) / 100.0)  # no step longer than 100


def test_march_away_from_the_surface_stops_at_max_distance():
    c = sphere_cfg()
    c.render[4] = 5000.0
    out = bind.march(c, [[0.0, 0.0, -1000.0]], [[0.0, 0.0, -1.0]], 64)
    travelled = np.linalg.norm(out[0, :3] - [0.0, 0.0, -1000.0])
    assert 5000.0 <= travelled < 5000.0 + 3072.0


def test_iteration_limit_mode_interior_and_escaped_rule():
    # bailout 2.5: the non-escaping ball is |z| <= 2.5 / 16, just outside the
    # sphere |z| = 2 / 16, so escaped points near it have a small raw distance
    c = sphere_cfg(mode=2, factor=0.1, bailout=2.5)
    inside = np.array([[0.0, 0.0, 100.0], [150.0, 0.0, 0.0]])  # |p| <= 2.5 ws / 16 = 160
    assert np.all(bind.map_distance(c, inside) == 0.0)
    p = np.array([[0.0, 170.0, 0.0]])
    raw = bind.kifs_sample(c, p)[0]
    assert raw[3] < 0 and 0 < raw[0] < np.linalg.norm(p - [0, 0, -1000]) * 0.1
    thr = np.float32(np.linalg.norm(np.array([0, 170, 1000.0]))) * np.float32(0.1)
    assert bind.map_distance(c, p)[0] == pytest.approx(1.01 * float(thr), rel=1e-6)
    c1 = sphere_cfg(mode=1, factor=0.1, bailout=2.5)
    assert bind.map_distance(c1, p)[0] == pytest.approx(raw[0], rel=1e-12)


@pytest.mark.parametrize("normal_mode", [0, 1])
def test_normals_on_the_sphere_are_radial(normal_mode):
    c = sphere_cfg(mode=0, thr=0.5)
    c.sdf_normal_mode = normal_mode
    rng = np.random.default_rng(3)
    u = rng.standard_normal((100, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    n = bind.normal(c, u * (R + 0.5))
    angle = np.arccos(np.clip(np.einsum("ij,ij->i", n, u), -1, 1))
    # the central stencil (step e = 0.5) is second order: angle O(e^2 / R^2);
    # the tetrahedral one is first order: O(h / R) with h = e / sqrt(3)
    bound = 1e-4 if normal_mode == 0 else 2 * 0.5 / math.sqrt(3) / R
    assert np.all(angle < bound), angle.max()


def test_slow_shading_normal_follows_the_iteration_count():
    # render[7] > 0: the 11^3 stencil over the iteration-count field.  The
    # escape iteration drops from 4 to 3 at |z| = 100 / 8 (world 12800), so
    # on an axis through that shell the normal is the axis, pointing outward.
    c = sphere_cfg(mode=0, thr=0.5)
    c.render[7] = 0.05
    c.camera_position[:] = [0.0, 0.0, -40000.0]
    for axis in range(3):
        p = np.zeros((1, 3))
        p[0, axis] = 12800.0
        n = bind.normal(c, p)[0]
        e = np.zeros(3)
        e[axis] = 1.0
        assert np.allclose(n, e, atol=1e-12)
    # deep inside the iteration-limit ball the count is constant: the
    # stencil has no direction and the reference returns +x
    assert np.allclose(bind.normal(c, [[0.0, 0.0, 500.0]])[0], [1.0, 0.0, 0.0])


def _rays(c, pixels):
    return bind.camera_ray(c, pixels)


def test_pixel_offset_and_perspective_projection():
    c = sphere_cfg()
    r = _rays(c, [[0, 31], [16, 7], [63, 0]])
    assert r[0, 0] == -1.0 and r[0, 1] == 0.5  # (0 / 64 - 0.5) * 2, (31 + 1) / 32 - 0.5
    assert r[1, 0] == -0.5 and r[1, 1] == -0.25
    f = 1 / math.tan(math.radians(30.0))
    for u, v, dx, dy, dz, seen in r:
        want = np.array([u, v, f]) / np.linalg.norm([u, v, f])
        assert seen == 1.0 and np.allclose([dx, dy, dz], want, atol=1e-12)


def test_fisheye_and_cut_fisheye():
    for proj in (1, 3):
        c = sphere_cfg()
        c.vset_values[107] = proj
        c.camera_fov = 200.0
        px = [[x, y] for x in range(0, 64, 3) for y in range(0, 32, 3)]
        r = _rays(c, px)
        fov = math.radians(200.0)
        for u, v, dx, dy, dz, seen in r:
            rad = math.hypot(u, v)
            cut = proj == 3 and rad * fov > math.pi / 2
            assert seen == (0.0 if cut else 1.0)
            if cut:
                continue
            # past pi off axis the ray has wrapped round the backward axis
            want = rad * fov if rad * fov <= math.pi else 2 * math.pi - rad * fov
            assert math.isclose(math.acos(max(-1.0, min(1.0, dz))), want, abs_tol=1e-7)
            if rad > 1e-6 and rad * fov < math.pi - 1e-6:
                assert math.isclose(math.atan2(dy, dx), math.atan2(v, u), abs_tol=1e-9)


def test_equirectangular_projection():
    c = sphere_cfg()
    c.vset_values[107] = 2
    c.camera_fov = 150.0
    fov = math.radians(150.0)
    aspect = 64 / 32
    for u, v, dx, dy, dz, seen in _rays(c, [[x, y] for x in range(0, 64, 5) for y in range(0, 32, 5)]):
        assert seen == 1.0
        assert math.isclose(math.atan2(dx, dz), u * (2 / aspect) * fov / 2, abs_tol=1e-9)
        assert math.isclose(math.asin(dy), v * fov / 2, abs_tol=1e-9)


def test_camera_rotation_is_the_shared_camera_rotation():
    """The Mandelbulber perspective ray is rotated like the standard camera:
    at screen points (0, 0) and (0, 0.25), reached exactly by both pixel
    conventions, the two cameras give the same rays."""
    def rotated(cfg):
        cfg.camera_yaw_pitch[:] = [0.4, -0.3]
        cfg.camera_roll = 0.2
        return cfg
    mb = rotated(sphere_cfg())
    std = rotated(sphere_cfg())
    std.sdf_id = ffi.SDF_MANDELBOX_FRACTAL
    # Mandelbulber 64 x 32: (32, 15) -> (0, 0), (32, 23) -> (0, 0.25)
    a = _rays(mb, [[32, 15], [32, 23]])
    assert a[0, 0] == 0 and a[0, 1] == 0 and a[1, 0] == 0 and a[1, 1] == 0.25
    # pixel centres: 63 x 31 (31, 15) -> (0, 0); 63 x 30 (31, 22) -> (0, 0.25)
    std.width, std.height = 63, 31
    b0 = _rays(std, [[31, 15]])[0]
    std.height = 30
    b1 = _rays(std, [[31, 22]])[0]
    assert b0[0] == 0 and b0[1] == 0 and b1[0] == 0 and b1[1] == 0.25
    assert np.allclose(a[0, 2:5], b0[2:5], atol=1e-14) and np.allclose(a[1, 2:5], b1[2:5], atol=1e-14)
    assert not np.allclose(a[0, 2:5], [0, 0, 1], atol=1e-3)


def test_focus_estimate_is_the_sphere_hit():
    c = sphere_cfg()
    f = bind.focus_estimate(c)
    assert 1000 - R - 2.0 < f < 1000 - R - 2.0 * 0.998


def test_diagnostic_hit_mask_and_depth_on_the_sphere():
    """Single-ray diagnostics of a Mandelbulber scene (SM:7304-7339): the
    Mandelbulber camera and march.  A pixel is a hit when its ray passes the
    sphere (plus the threshold band) and a miss when it passes clear of it;
    the depth is 1 - t / max_distance with t the distance to the band."""
    c = sphere_cfg(mode=0, thr=2.0)
    c.width, c.height, c.camera_fov = 48, 32, 20.0
    md = 2000.0
    hm = bind.diagnostic(c, ffi.DIAGNOSTIC_MODES["hit_mask"], max_distance=md)[::-1, :, 0]  # PNG rows are flipped
    dp = bind.diagnostic(c, ffi.DIAGNOSTIC_MODES["depth"], max_distance=md)[::-1, :, 0]
    px = [[x, y] for y in range(32) for x in range(48)]
    d = bind.camera_ray(c, px)[:, 2:5]
    cam = np.array([0.0, 0.0, -1000.0])
    along = -(d @ cam)  # distance along each ray to its closest approach to the origin
    miss_dist = np.linalg.norm(cam + d * along[:, None], axis=1)
    inside, outside = miss_dist < R + 2.0 * LIMIT - 1e-6, miss_dist > R + 2.0 + 1e-6
    assert inside.sum() > 100 and outside.sum() > 100
    hm = hm.reshape(-1)
    assert np.all(hm[inside] == 255) and np.all(hm[outside] == 0)
    # entry distance into the sphere of radius R + h for h in the band
    def t_at(radius):
        return along[inside] - np.sqrt(np.maximum(radius ** 2 - miss_dist[inside] ** 2, 0.0))
    lo, hi = t_at(R + 2.0), t_at(R + 2.0 * LIMIT)
    v = dp.reshape(-1)[inside].astype(float)
    assert np.all(v >= np.floor((1 - hi / md) * 255) - 1) and np.all(v <= np.ceil((1 - lo / md) * 255) + 1)
