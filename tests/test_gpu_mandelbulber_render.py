"""Rendering imported Mandelbulber scenes on the GPU (the Kaleidoscopic IFS
field, formula 10) against the oracle: the reference's Mandelbulber camera,
distance, march and normal branches (Shaders.metal:473-511, 975-995,
3627-3751, 3871-3960) inside renderPath (5926-6026), on the
one-path-per-lane kernel (and the persistent kernel as its A/B twin).

Tolerances, as for the other SDF scenes:
  closed-form sphere   every GPU hit ends in the refinement band above the
                       sphere (threshold * (1 - 0.001 detail), threshold),
                       within FP32 rounding of the position;
  per-ray (ifs-20)     hit masks agree on >= 99% of pixels; hit distances
                       within 1e-4 relative for at least as many rays as the
                       FP32 twin of the oracle manages (less 1%), each
                       measured against the FP64 oracle;
  images               RMSE to the FP64 image at most 1.25x the FP32 twin's
                       plus 2e-3 (chaotic divergence of FP32 paths);
  work counters        within 2% of the oracle's (distance evaluations,
                       march and shadow steps, normals).
"""
from pathlib import Path

import numpy as np
import pytest

from fractalpathtracer_amd import api, ffi, scenes
from kifs_sphere import LIMIT, R, sphere_cfg
from oracle import bind

pytestmark = pytest.mark.gpu

IFS20 = Path(__file__).resolve().parent / "golden" / "mandelbulber" / "ifs-20.fract"


def _ifs20(w=64, h=36, spp=2):
    c, _ = scenes.load_scene(IFS20)
    c.width, c.height, c.samples = w, h, spp
    return c


@pytest.mark.parametrize("mode,thr", [(0, 2.0), (1, 0.002)])
def test_sphere_hits_end_in_the_refinement_band(mode, thr):
    c = sphere_cfg(mode=mode, thr=thr, factor=thr)
    c.camera_fov = 20.0
    h = api.trace_primary(c)
    r = bind.trace_primary(c, 64)
    hit = h["hit"] == 1
    assert np.array_equal(hit, r[..., 7] == 1) and hit.sum() > 200
    p = h["position"][hit].astype(np.float64)
    height = np.linalg.norm(p, axis=-1) - R
    if mode == 0:
        lo, hi = np.full(len(p), thr * LIMIT), np.full(len(p), thr)
    else:
        # the threshold of the last march step, at that point's distance from
        # the camera (the refinement moves the point by less than it)
        dist = np.linalg.norm(p - [0.0, 0.0, -1000.0], axis=-1)
        lo, hi = dist * thr * LIMIT * (1 - 1e-2), dist * thr * (1 + 1e-2)
    slack = 1e-3  # FP32 rounding of positions ~1000 world units from the origin
    print(mode, "height", height.min(), height.max())
    assert np.all(height > lo - slack) and np.all(height < hi + slack)
    n = h["normal"][hit]
    assert np.all(np.einsum("ij,ij->i", n, p / np.linalg.norm(p, axis=-1, keepdims=True)) > 1 - 1e-4)


def test_ifs20_primary_rays_match_oracle():
    c = _ifs20(96, 54)
    h = api.trace_primary(c)
    r = bind.trace_primary(c, 64)
    tw = bind.trace_primary(c, 32)
    hit_g, hit_r, hit_t = h["hit"] == 1, r[..., 7] == 1, tw[..., 7] == 1
    print("hit agreement gpu %.4f twin %.4f" % (np.mean(hit_g == hit_r), np.mean(hit_t == hit_r)))
    assert np.mean(hit_g == hit_r) >= 0.99
    both = hit_g & hit_r & hit_t
    t_ref = r[..., 0][both]
    err_g = np.abs(h["t"][both] - t_ref) / t_ref
    err_t = np.abs(tw[..., 0][both] - t_ref) / t_ref
    fg, ft = np.mean(err_g <= 1e-4), np.mean(err_t <= 1e-4)
    print("t within 1e-4: gpu %.4f twin %.4f; median gpu %.2e twin %.2e" % (fg, ft, np.median(err_g),
                                                                          np.median(err_t)))
    assert fg >= ft - 0.01


@pytest.mark.parametrize("case", ["ifs20", "ifs20_dof", "fisheye_cut", "equirect"])
def test_image_parity_against_fp64_and_fp32_twin(case):
    c = _ifs20(64, 36, 2)
    if case == "ifs20_dof":
        c.camera_dof, c.focus_distance = 0.05, 0.0  # the focus prepass estimates it
    elif case == "fisheye_cut":
        c.vset_values[107], c.camera_fov = 3.0, 200.0
    elif case == "equirect":
        c.vset_values[107], c.camera_fov = 2.0, 150.0
    img = api.render(c)
    ref = bind.render(c, 64)
    twin = bind.render(c, 32)
    rmse = float(np.sqrt(np.mean((img[..., :3] - ref[..., :3]) ** 2)))
    rmse_twin = float(np.sqrt(np.mean((twin[..., :3] - ref[..., :3]) ** 2)))
    print(case, "rmse gpu %.3e twin %.3e mean %.3f" % (rmse, rmse_twin, float(img[..., :3].mean())))
    assert np.all(np.isfinite(img)) and img[..., :3].mean() > 0.02
    assert rmse <= 1.25 * rmse_twin + 2e-3
    if case == "fisheye_cut":
        # pixels whose every jittered camera sample (jitter <= 0.3 / max(W, H)
        # in screen units) lies beyond 90 degrees off axis see nothing
        uv = bind.camera_ray(c, [[x, y] for y in range(c.height) for x in range(c.width)])[:, :2]
        edge = 0.5 * np.pi / np.radians(c.camera_fov)
        outside = (np.hypot(uv[:, 0], uv[:, 1]) - edge).reshape(c.height, c.width) > 2 * 0.3 / c.width
        assert outside.any()
        assert np.all(img[..., :3][outside] == 0) and np.all(ref[..., :3][outside] == 0)


def test_profile_counters_match_oracle_work():
    c = _ifs20(48, 27, 1)
    res = api.render_full(c, profile=True)
    bind.render(c, 64)
    k = bind.counters()
    p = res.profile
    assert p["pixels"] == 48 * 27
    for key in ("distance_evals", "primary_steps", "shadow_steps", "normal_evals"):
        print(key, p[key], k[key])
        assert abs(p[key] - k[key]) <= 0.02 * k[key], (key, p[key], k[key])


def test_persistent_and_one_path_kernels_agree():
    """Imported scenes render on the persistent kernel by default (8 waves per
    SIMD, the faster one for them) and on the one-path-per-lane kernel when
    a plan asks for it: two independent implementations of the same path,
    whose images must be bit-identical (cut fish-eye too)."""
    for proj, fov in ((0.0, None), (3.0, 200.0)):
        c = _ifs20(48, 27, 3)
        c.vset_values[107] = proj
        if fov:
            c.camera_fov = fov
        imgs = []
        for kernel in (1, 0):
            with api.Plan(c) as plan:
                plan.tune(kernel=kernel)
                buf = api.DeviceBuffer(48 * 27 * 4)
                plan.render_into(buf.address)
                imgs.append(buf.numpy().reshape(27, 48, 4))
        assert np.array_equal(imgs[0], imgs[1]), proj
        assert np.array_equal(imgs[0], api.render(c))


def test_sample_runs_on_imported_scenes():
    """The persistent kernel's Mandelbulber instance in sample runs (an
    8-aligned frame, 20 and 37 samples split into runs) renders the same
    image as the one-path-per-lane kernel, chunked launches included."""
    for spp, chunked in ((20, False), (37, True)):
        c = _ifs20(40, 32, spp)
        if chunked:
            c.sdf_accumulation_mode = ffi.SDF_ACCUMULATION_CHUNKED
            c.sdf_chunk_samples = 17
        imgs = []
        for kernel, runs in ((0, 1), (1, 8), (1, 3)):
            with api.Plan(c) as plan:
                plan.tune(kernel=kernel)
                plan.tune_stash(6, runs)
                buf = api.DeviceBuffer(40 * 32 * 4)
                plan.render_into(buf.address)
                imgs.append(buf.numpy().copy())
        assert np.array_equal(imgs[0], imgs[1]), spp
        assert np.array_equal(imgs[0], imgs[2]), spp
