"""The GPU path tracer's estimator branches against closed forms and the
FP64 oracle (renderPath, Shaders.metal:5926-6026): emission, the light-sum x
throughput composition, the bounce caps, Russian roulette, refraction,
Fresnel reflection, total internal reflection and the translucency
threshold of the march (SM:3777).  Both SDF kernels are run: the
persistent path-regenerating kernel (plans' default) and the one-path-per-
lane kernel."""
import math

import numpy as np
import pytest

from fractalpathtracer_amd import api, ffi, scenes
from oracle import bind
from tests import estimator_scenes as es
from tests.test_estimator_oracle import _cavity_closed_form, rr_expectation, rr_sigma

pytestmark = pytest.mark.gpu


def _plan_render(cfg, kernel):
    with api.Plan(cfg) as plan:
        plan.tune(kernel=kernel)
        buf = api.DeviceBuffer(cfg.width * cfg.height * 4)
        plan.render_into(buf.address)
        return buf.numpy().reshape(cfg.height, cfg.width, 4)


@pytest.mark.parametrize("kernel", [1, 0])
@pytest.mark.parametrize("bounces,cap,albedo,e", [
    (3, 0, (0.9, 0.6, 0.3), 1.5),
    (1, 0, (0.8, 0.4, 0.2), 2.0),
    (12, 0, (0.95, 0.7, 0.5), 1.0),
    (6, 4, (0.9, 0.6, 0.3), 1.5),
    (3, 0, (0.9, 0.6, 0.3), 10.0),
])
def test_emissive_cavity_closed_form_on_the_gpu(kernel, bounces, cap, albedo, e):
    """Every sample of every pixel is N a e a^N (clamped at 8)."""
    cfg = es.cavity_cfg(bounces=bounces, bounce_cap=cap, albedo=albedo, emission=e, spp=3)
    img = _plan_render(cfg, kernel)
    n = min(bounces, 8) if cap == 0 else min(bounces, 8, cap)
    want = _cavity_closed_form(n, albedo, e)
    err = np.abs(img[..., :3] - want) / np.maximum(want, 1e-30)
    print(f"kernel {kernel} N={n}: max relative error {err.max():.2e}")
    assert err.max() <= 2e-5


@pytest.mark.parametrize("kernel", [1, 0])
def test_russian_roulette_expectation_on_the_gpu(kernel):
    a, e = 0.5, 1.0
    cfg = es.cavity_cfg(bounces=4, albedo=(a, a, a), emission=e, width=256, height=256, spp=8, rr=True,
                        rr_start=2.0, rr_min=0.05)
    img = _plan_render(cfg, kernel)
    n = cfg.width * cfg.height * cfg.samples
    mean = float(img[..., 0].mean())
    print(f"kernel {kernel}: mean {mean:.5f} expectation {rr_expectation(a, e):.5f}")
    assert abs(mean - rr_expectation(a, e)) < 5 * rr_sigma(a, e) / math.sqrt(n)


def test_russian_roulette_paths_match_the_oracle():
    """Same RNG keys, same survival decisions: the GPU image equals the FP64
    oracle's in the cavity, where every march ends on the same wall."""
    cfg = es.cavity_cfg(bounces=5, albedo=(0.7, 0.5, 0.3), emission=1.2, width=48, height=40, spp=4, rr=True,
                        rr_start=2.0, rr_min=0.2)
    img = _plan_render(cfg, 1)
    ref = bind.render(cfg, 64)
    close = np.isclose(img[..., :3], ref[..., :3], rtol=2e-5, atol=1e-6).all(axis=-1)
    print(f"pixels equal to the oracle within 2e-5: {close.mean():.4f}")
    assert close.mean() >= 0.999


@pytest.mark.parametrize("transl", [0.0, 0.3])
@pytest.mark.parametrize("min_distance", [0.01, 1e-6])
def test_translucent_threshold_step_counts_on_the_gpu(transl, min_distance):
    """Centre ray onto a unit sphere from distance 3: 3/5 steps with opaque
    thresholds 0.01/1e-6, 4 with a translucent material (the closed-form step
    sequence of the oracle pin)."""
    cfg = es.cavity_cfg(width=33, height=33)
    scenes.set_program(cfg, [(ffi.OP_SPHERE, 0, [1.0, 0, 0, 0])])
    cfg.camera_position[:] = [0.0, 0.0, -3.0]
    cfg.camera_yaw_pitch[:] = [0.0, 0.0]
    cfg.program_material[5] = transl
    cfg.render[3] = min_distance
    hits = api.trace_primary(cfg)
    want = {(0.0, 0.01): 3, (0.0, 1e-6): 5}.get((transl, min_distance), 4)
    centre = int(hits["steps"][16, 16])
    print(f"translucency {transl} min_distance {min_distance}: {centre} steps")
    assert centre == want
    ref = bind.trace_primary(cfg, 64)
    # Grazing rays at the silhouette may round differently in FP32, most of
    # all with the 1e-6 threshold (98.5% identical measured).
    same = float(np.mean(hits["steps"] == ref[..., 8].astype(np.uint32)))
    print(f"step counts identical to FP64 on {same:.4f} of the rays")
    assert same >= 0.97


def test_glass_scene_primary_rays_against_fp64():
    cfg = es.glass_cfg(width=96, height=72)
    hits = api.trace_primary(cfg)
    ref = bind.trace_primary(cfg, 64)
    both = (hits["hit"] == 1) & (ref[..., 7] == 1)
    assert both.mean() > 0.5
    assert np.mean(hits["hit"] == ref[..., 7]) >= 0.995
    rel = np.abs(hits["t"][both] - ref[..., 0][both]) / ref[..., 0][both]
    assert np.mean(rel <= 1e-4) >= 0.995
    same_steps = np.mean(hits["steps"][both] == ref[..., 8][both].astype(np.uint32))
    print(f"glass primary: t within 1e-4 {np.mean(rel <= 1e-4):.4f}, same step count {same_steps:.4f}")
    assert same_steps >= 0.98


@pytest.mark.parametrize("transl", [0.7, 1.0])
def test_glass_scene_image_parity_against_fp64(transl):
    """Refraction, Fresnel, total internal reflection, emission and the
    translucency threshold all run (translucency 1: every scatter is on the
    refractive branch)."""
    cfg = es.glass_cfg(spp=8, translucency=transl)
    img = _plan_render(cfg, 1)
    one_path = _plan_render(cfg, 0)
    assert np.array_equal(img, one_path)
    ref = bind.render(cfg, 64)
    twin = bind.render(cfg, 32)
    rmse = float(np.sqrt(np.mean((img[..., :3] - ref[..., :3]) ** 2)))
    rmse_twin = float(np.sqrt(np.mean((twin[..., :3] - ref[..., :3]) ** 2)))
    bias = float(np.mean(img[..., :3] - ref[..., :3]))
    print(f"glass translucency {transl}: rmse gpu {rmse:.3e} twin {rmse_twin:.3e} mean diff {bias:.2e}")
    assert np.all(np.isfinite(img))
    assert rmse <= 1.25 * rmse_twin + 2e-3
