"""Diagnostic images (fpt_metal_diagnostic_render / fpt_hip_diagnostic_render,
Shaders.metal:7282-7442) and the present pass (PNG) against the oracle.

8-bit outputs are compared per channel: the share of values within one step
of the oracle's.  Single-ray modes on a smooth scene must agree almost
everywhere; path modes on a fractal can take a different branch where the
FP32 GPU and the FP64 oracle land on opposite sides of a hit/miss or scatter
decision, which changes a whole pixel."""
from pathlib import Path

import numpy as np
import pytest

from fractalpathtracer_amd import api, ffi, scenes
from oracle import bind

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
MODES = ffi.DIAGNOSTIC_MODES
SINGLE_RAY = ["depth", "normal", "material", "hit_mask", "path_direct", "path_environment", "path_throughput",
              "diffuse_normal", "mandel_color_index", "mandel_palette_position"]
PATH_MODES = ["path_final", "sdf_primary_steps", "sdf_shadow_steps", "sdf_normal_evals", "sdf_bounces",
              "sdf_bounce_contribution"]


def _close(a, b):
    return float(np.mean(np.abs(a.astype(int) - b.astype(int)) <= 1))


def _sphere(w=48, h=32):
    c = scenes.apply_scene_json(scenes.default_config(), GOLDEN / "scenes" / "Exact_Sphere.json")
    c.width, c.height, c.samples = w, h, 2
    return c


def _bulb(w=48, h=36):
    c = scenes.mandelbulb_scene(w, h, spp=2, bounces=2)
    c.world[4] = 0.3   # some environment light so environment/path modes are not black
    c.world[0] = 3.0
    return c


@pytest.mark.parametrize("mode", SINGLE_RAY + PATH_MODES)
def test_sdf_diagnostic_sphere_matches_oracle(mode):
    c = _sphere()
    img, ms = api.diagnostic(c, mode, max_distance=8.0, bounce=1)
    ref = bind.diagnostic(c, MODES[mode], bounce=1, max_distance=8.0)
    frac = _close(img, ref)
    print(mode, f"within 1: {frac:.4f}  ms {ms:.3f}")
    assert np.all(img[..., 3] == 255)
    assert frac >= (0.995 if mode in SINGLE_RAY else 0.98)


@pytest.mark.parametrize("mode", SINGLE_RAY + PATH_MODES)
def test_sdf_diagnostic_mandelbulb_matches_oracle(mode):
    """On the fractal the bar is the oracle's own FP32 twin: the GPU may not
    be further from the FP64 oracle than plain FP32 arithmetic of the same
    algorithm is (mean 8-bit error within 1.25x the twin's + 0.5 steps)."""
    c = _bulb()
    img, _ = api.diagnostic(c, mode, bounce=0)
    ref = bind.diagnostic(c, MODES[mode], bounce=0)
    twin = bind.diagnostic(c, MODES[mode], bounce=0, precision=32)
    mad = float(np.mean(np.abs(img.astype(int) - ref.astype(int))))
    mad_twin = float(np.mean(np.abs(twin.astype(int) - ref.astype(int))))
    frac = _close(img, ref)
    print(mode, f"within 1: {frac:.4f} twin {_close(twin, ref):.4f}; mean |d| {mad:.3f} twin {mad_twin:.3f}")
    assert mad <= 1.25 * mad_twin + 0.5
    if mode in SINGLE_RAY:
        assert frac >= 0.97


@pytest.mark.parametrize("mode", ["depth", "normal", "material", "hit_mask", "path_direct", "path_environment",
                                  "path_throughput", "diffuse_normal", "sdf_primary_steps", "path_final"])
def test_voxel_diagnostic_matches_oracle(mode):
    c = scenes.voxel_mandelbox_scene(40, 30, spp=1, resolution=64)
    c.world[0], c.world[4] = 3.0, 0.3
    img, _ = api.diagnostic(c, mode)
    ref = bind.diagnostic(c, MODES[mode])
    frac = _close(img, ref)
    print(mode, f"within 1: {frac:.4f}")
    assert frac >= 0.99


def test_diagnostic_png_equals_buffer(tmp_path):
    c = _bulb(40, 24)
    buf, _ = api.diagnostic(c, "normal")
    png, ms = api.diagnostic(c, "normal", output_png=str(tmp_path / "d.png"))
    assert ms > 0 and np.array_equal(buf, png)


def test_diagnostic_errors(tmp_path):
    c = _bulb(8, 8)
    lib = ffi.load_library()
    err = ffi.error_buffer()
    import ctypes as C
    dg = ffi.FptDiagnosticConfig(3, 0, 0.0, 0.0)
    assert lib.fpt_metal_diagnostic_render(None, None, None, 0, C.byref(c), C.byref(dg), None, err, len(err)) == 1
    assert b"output path" in err.value
    with pytest.raises(ffi.FptError, match="unknown diagnostic mode"):
        api.diagnostic(c, 99)


def test_present_png_matches_oracle_present(tmp_path):
    """The PNG of fpt_metal_render = oracle present of the same call's linear
    buffer (tone map, exposure, contrast, saturation, aberration, bloom,
    flip)."""
    c = scenes.mandelbulb_scene(64, 48, 2, bounces=1)
    for i, v in enumerate([0.0, 1.3, 0.02, 0.9, 1.1, 2.0, 0.4]):
        c.post[i] = v
    res = api.render_full(c, output_png=str(tmp_path / "p.png"))
    png = api.read_png(str(tmp_path / "p.png"))
    ref = bind.present(c, res.linear.astype(np.float32))
    frac = _close(png, ref)
    print("present within 1:", frac, "exact:", float(np.mean(png == ref)))
    assert frac >= 0.999
    for tm in (1.0, 3.0, 4.0):
        c.post[0] = tm
        res = api.render_full(c, output_png=str(tmp_path / "q.png"))
        assert _close(api.read_png(str(tmp_path / "q.png")), bind.present(c, res.linear.astype(np.float32))) >= 0.999


def _kifs_sphere():
    """The closed-form KIFS sphere with a fixed threshold of 1e-3 (normal
    step 0.1): the path tracer's 0.001 surface offset then lifts secondary
    rays clear of the hit band, so paths are well conditioned."""
    from kifs_sphere import sphere_cfg
    c = sphere_cfg(mode=0, thr=1e-3)
    c.vset_values[114] = 100.0
    c.width, c.height, c.samples = 48, 32, 2
    c.camera_fov = 20.0
    c.world[0], c.world[4] = 3.0, 0.3
    c.sun[:] = [1.0, 135.0, 35.26, 1.0, 0.0]
    return c


def _ifs20():
    c, _ = scenes.load_scene(GOLDEN / "mandelbulber" / "ifs-20.fract")
    c.width, c.height, c.samples = 48, 27, 2
    return c


@pytest.mark.parametrize("mode", SINGLE_RAY + PATH_MODES)
def test_sdf_diagnostic_kifs_sphere_matches_oracle(mode):
    """Imported-scene diagnostics (Mandelbulber camera, march, normals) on
    the KIFS whose field is an exact sphere distance.  Single-ray modes: the
    smooth-scene bar.  Path modes: the FP32 twin's bar, since a few paths
    graze the silhouette (the twin itself agrees with FP64 on 98% of
    path_final values)."""
    c = _kifs_sphere()
    img, _ = api.diagnostic(c, mode, max_distance=2000.0, bounce=1)
    ref = bind.diagnostic(c, MODES[mode], bounce=1, max_distance=2000.0)
    twin = bind.diagnostic(c, MODES[mode], bounce=1, max_distance=2000.0, precision=32)
    frac = _close(img, ref)
    mad = float(np.mean(np.abs(img.astype(int) - ref.astype(int))))
    mad_twin = float(np.mean(np.abs(twin.astype(int) - ref.astype(int))))
    print(mode, f"within 1: {frac:.4f} twin {_close(twin, ref):.4f}; mean |d| {mad:.3f} twin {mad_twin:.3f}")
    assert np.all(img[..., 3] == 255)
    if mode in SINGLE_RAY:
        assert frac >= 0.995
    else:
        assert mad <= 1.25 * mad_twin + 0.5


@pytest.mark.parametrize("mode", SINGLE_RAY + PATH_MODES)
def test_sdf_diagnostic_ifs20_matches_oracle(mode):
    """The imported ifs-20 scene: the fractal bar (mean 8-bit error within
    1.25x the FP32 twin's + 0.5 steps)."""
    c = _ifs20()
    img, _ = api.diagnostic(c, mode, bounce=0)
    ref = bind.diagnostic(c, MODES[mode], bounce=0)
    twin = bind.diagnostic(c, MODES[mode], bounce=0, precision=32)
    mad = float(np.mean(np.abs(img.astype(int) - ref.astype(int))))
    mad_twin = float(np.mean(np.abs(twin.astype(int) - ref.astype(int))))
    print(mode, f"within 1: {_close(img, ref):.4f} twin {_close(twin, ref):.4f}; mean |d| {mad:.3f} twin {mad_twin:.3f}")
    assert mad <= 1.25 * mad_twin + 0.5
