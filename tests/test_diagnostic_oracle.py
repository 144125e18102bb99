"""Pins of the oracle's present and diagnostic images (no GPU): tone-map
closed forms, highlight/chromatic/flip behaviour on constructed buffers, and
the Exact_Sphere fixture's closed-form hit mask, depth, normal, diffuse
normal and bounce-count images (Shaders.metal:6778-6985, 7282-7373,
7462-7507)."""
from pathlib import Path

import numpy as np

from fractalpathtracer_amd import ffi, scenes
from oracle import bind

GOLDEN = Path(__file__).resolve().parent / "golden"
MODES = ffi.DIAGNOSTIC_MODES


def _flat(w, h, rgb):
    a = np.zeros((h, w, 4), np.float32)
    a[..., :3] = rgb
    a[..., 3] = 1.0
    return a


def _cfg(w=8, h=4, post=(0.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0)):
    c = scenes.default_config()
    c.width, c.height = w, h
    for i, v in enumerate(post):
        c.post[i] = v
    return c


def _srgb(v):
    v = np.asarray(v, np.float64)
    return np.where(v <= 0.0031308, v * 12.92, 1.055 * np.power(np.maximum(v, 0), 1 / 2.4) - 0.055)


def test_present_srgb_tone_map_closed_form_and_truncation():
    vals = np.array([0.0, 0.001, 0.0031308, 0.01, 0.18, 0.5, 1.0, 3.0], np.float64)
    c = _cfg(w=len(vals), h=1)
    acc = _flat(len(vals), 1, 0.0)
    acc[0, :, 0] = vals
    acc[0, :, 1] = vals
    acc[0, :, 2] = vals
    out = bind.present(c, acc)
    want = np.floor(np.clip(_srgb(vals), 0, 1) * 255).astype(np.uint8)
    # truncation of x*255 (uchar conversion): srgb(1) = 1.055 - 0.055 lands a
    # hair under 1 -> 254; values above 1 clamp to exactly 1 -> 255
    assert np.array_equal(out[0, :, 0], want) and np.all(out[..., 3] == 255)
    assert out[0, 6, 0] == 254 and out[0, 7, 0] == 255 and out[0, 0, 0] == 0


def test_present_other_tone_maps_fixed_points():
    # ACES fitted curve (Narkowicz): f(0) = 0, f(large) -> 1 (clamped);
    # power 1/2.4 maps 1 to 1; PBR neutral leaves values under the knee alone
    # except for the toe offset (x - 6.25 x^2 for x < 0.08).
    for tm, v, want in [(3.0, 0.0, 0), (3.0, 1000.0, 255), (1.0, 1.0, 255), (1.0, 0.5 ** 2.4, 127)]:
        c = _cfg(w=1, h=1, post=(tm, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0))
        out = bind.present(c, _flat(1, 1, v))
        assert out[0, 0, 0] == want, (tm, v, out[0, 0])
    c = _cfg(w=1, h=1, post=(4.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0))
    x = 0.05
    out = bind.present(c, _flat(1, 1, x))
    assert out[0, 0, 0] == int((x - (x - 6.25 * x * x)) * 255)  # grey: c - off = 6.25 x^2


def test_present_exposure_contrast_saturation():
    # linear tone map (mode 5 = identity), exposure 2, brightness 0.1,
    # saturation 0 (luminance), contrast 1
    c = _cfg(w=1, h=1, post=(5.0, 2.0, 0.1, 0.0, 1.0, 0.0, 0.0))
    acc = _flat(1, 1, 0.0)
    acc[0, 0, :3] = [0.1, 0.2, 0.3]
    out = bind.present(c, acc)
    rgb = np.array([0.1, 0.2, 0.3]) * 2 + 0.1
    lum = float(rgb @ [0.2126, 0.7152, 0.0722])
    assert np.all(out[0, 0, :3] == int(lum * 255))


def test_present_highlight_bloom_and_flip():
    w, h = 5, 3
    k = 0.5
    c = _cfg(w, h, post=(5.0, 1.0, 0.0, 1.0, 1.0, 0.0, k))
    v = 0.3
    out = bind.present(c, _flat(w, h, v))
    # constant image: bloom = k (max(v - 0.5, 0) + 4 * 0.25 v)
    assert np.all(out[..., 0] == int((v + k * (max(v - 0.5, 0) + v)) * 255))
    # flip: scene row y lands at output row h-1-y
    c2 = _cfg(w, h, post=(5.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0))
    acc = _flat(w, h, 0.0)
    for y in range(h):
        acc[y, :, 0] = 0.25 * (y + 1)
    out2 = bind.present(c2, acc)
    for y in range(h):
        assert np.all(out2[h - 1 - y, :, 0] == int(0.25 * (y + 1) * 255))


def test_present_chromatic_aberration_is_identity_on_constant_images():
    c = _cfg(6, 4, post=(5.0, 1.0, 0.0, 1.0, 1.0, 3.0, 0.0))
    out = bind.present(c, _flat(6, 4, [0.2, 0.4, 0.6]))
    assert np.all(out[..., 0] == int(np.float32(0.2) * 255)) and np.all(out[..., 2] == int(np.float32(0.6) * 255))


# --- diagnostics on the Exact_Sphere fixture -----------------------------------

def _sphere(w=48, h=32):
    c = scenes.apply_scene_json(scenes.default_config(), GOLDEN / "scenes" / "Exact_Sphere.json")
    c.width, c.height, c.samples = w, h, 2
    return c


def _sphere_rays(c):
    """Closed-form camera rays (Shaders.metal:6847-6855, 460-471) and their
    unit-sphere hits: (hit, t, normal) per pixel in PNG row order."""
    w, h = c.width, c.height
    ys, xs = np.mgrid[0:h, 0:w]
    u = ((xs + 0.5) / w - 0.5) * (w / h)
    v = (ys + 0.5) / h - 0.5
    f = 1.0 / np.tan(np.radians(c.camera_fov) / 2)
    d = np.stack([u, v, np.full_like(u, f)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)   # yaw = pitch = roll = 0
    o = np.array(c.camera_position[:3], np.float64)
    b = d @ o
    disc = b * b - (o @ o - 1.0)
    hit = disc > 0
    t = np.where(hit, -b - np.sqrt(np.maximum(disc, 0)), np.inf)
    n = o + d * t[..., None]
    flip = lambda a: a[::-1]  # noqa: E731  (scene row y -> PNG row h-1-y)
    return flip(hit), flip(t), flip(n), flip(np.abs(disc))


def test_sphere_hit_mask_depth_normal_closed_forms():
    c = _sphere()
    hit, t, n, disc = _sphere_rays(c)
    safe = disc > 0.05  # one pixel away from the silhouette
    mask = bind.diagnostic(c, MODES["hit_mask"])
    assert np.all((mask[..., 0] == 255)[safe] == hit[safe])
    depth = bind.diagnostic(c, MODES["depth"], max_distance=8.0)
    # the march stops within max(min_distance, 5e-4) of the surface
    want = np.floor((1 - np.clip(t / 8.0, 0, 1)) * 255)
    sel = safe & hit
    assert np.all(np.abs(depth[..., 0][sel] - want[sel]) <= 1)
    assert np.all(depth[..., 0][safe & ~hit] == 0)
    nimg = bind.diagnostic(c, MODES["normal"])
    want_n = np.floor(np.clip(n * 0.5 + 0.5, 0, 1) * 255)
    assert np.all(np.abs(nimg[..., :3][sel].astype(int) - want_n[sel]) <= 1)
    dn = bind.diagnostic(c, MODES["diffuse_normal"])
    want_d = np.floor((0.12 + 0.88 * np.maximum(-n[..., 2], 0)) * 255)  # camera looks along +z
    assert np.all(np.abs(dn[..., 0][sel].astype(int) - want_d[sel]) <= 1)


def test_sphere_bounce_count_closed_form():
    """A convex scene: every camera hit scatters away and escapes, so each
    sample of a hit pixel has exactly one bounce (value 1 / render[0])."""
    c = _sphere()
    hit, _, _, disc = _sphere_rays(c)
    img = bind.diagnostic(c, MODES["sdf_bounces"])
    safe = disc > 0.05
    want = np.where(hit, int(1.0 / max(c.render[0], 1.0) * 255), 0)
    assert np.all(img[..., 0][safe] == want[safe])
    # normal evaluations = bounces, scaled by 2 * render[0]
    ne = bind.diagnostic(c, MODES["sdf_normal_evals"])
    want_ne = np.where(hit, int(1.0 / (2 * c.render[0]) * 255), 0)
    assert np.all(ne[..., 0][safe] == want_ne[safe])


def test_path_final_diagnostic_is_post_processed_render_mean():
    """PATH_FINAL = post-process of the mean of min(max(spp, 16), 64) path
    samples: the same as presenting a 16-sample oracle render without bloom
    or aberration."""
    c = _sphere(24, 16)
    c.samples = 2
    img = bind.diagnostic(c, MODES["path_final"], precision=64)
    c16 = scenes.clone(c)
    c16.samples = 16
    acc = bind.render(c16, 64).astype(np.float32)
    c16.post[5] = 0.0
    c16.post[6] = 0.0
    want = bind.present(c16, acc)
    # running mean (render) vs sum / n (diagnostic) differ in the last bits
    assert np.mean(np.abs(img.astype(int) - want.astype(int)) <= 1) > 0.99
