"""GPU parity of the SDF path (HIP, through the C ABI) against the FP64 oracle.

Tolerances (FP32 GPU vs FP64 CPU on chaotic fractals), following the
reference's own qualification style for GPU-f32 vs CPU-f64 fields
(src/main.rs:3297-3361: per-sample tolerance, failure fraction, p99):
  per-ray hit t and position   relative error <= 1e-4 for >= 99.5% of rays
                               hit by both, p99 <= 1e-5; hit masks agree on
                               >= 99.9% of pixels;
  normals                      the GPU normal against the FP64 normal at the
                               GPU's own hit point, within max(1e-4, 4 s) for
                               >= 98.5% of rays, where s is how far the FP64
                               normal itself moves when the point moves by one
                               float32 ulp along an axis (the 12-iteration DE's
                               gradient at the surface is that ill-conditioned:
                               median s ~ 1.4e-4 on the Mandelbulb);
  field samples                |d - d64| <= 5e-5 (1 + |d64|) for every point;
  images                       the GPU image's RMSE to the FP64 image is no
                               more than 1.25x that of the FP32 CPU twin (the
                               same algorithm in FP32) plus 2e-3: chaotic
                               divergence of FP32 paths is intrinsic.
"""
import numpy as np
import pytest

from fractalpathtracer_amd import api, ffi, scenes
from oracle import bind

pytestmark = pytest.mark.gpu

GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def _scene_json(name, w, h, spp=1):
    c = scenes.apply_scene_json(scenes.default_config(), GOLDEN / "scenes" / f"{name}.json")
    c.width, c.height, c.samples = w, h, spp
    return c


CASES = {
    "mandelbulb_p8": lambda: scenes.mandelbulb_scene(96, 96, 1),
    "typed_program": lambda: scenes.typed_program_scene(128, 72, 1),
    "exact_csg": lambda: _scene_json("Exact_CSG", 96, 54),
    "fold_program": lambda: _scene_json("Typed_Program_Fold", 96, 72),
}


def test_device_is_gfx950():
    name = api.device_name()
    print(name)
    assert "gfx950" in name


@pytest.mark.parametrize("case", list(CASES))
def test_primary_rays_match_fp64_oracle(case):
    cfg = CASES[case]()
    h = api.trace_primary(cfg)
    r = bind.trace_primary(cfg, 64)
    hit_g, hit_r = h["hit"] == 1, r[..., 7] == 1
    assert np.mean(hit_g == hit_r) >= 0.999
    both = hit_g & hit_r
    assert both.sum() > 100
    t_err = np.abs(h["t"][both] - r[..., 0][both]) / r[..., 0][both]
    p_err = np.linalg.norm(h["position"][both] - r[..., 1:4][both], axis=-1) / np.linalg.norm(
        r[..., 1:4][both], axis=-1).clip(1e-6)
    print(case, "t: frac<=1e-4 %.5f p99 %.2e" % (np.mean(t_err <= 1e-4), np.percentile(t_err, 99)))
    assert np.mean(t_err <= 1e-4) >= 0.995 and np.percentile(t_err, 99) <= 1e-5
    assert np.mean(p_err <= 1e-4) >= 0.995
    # Normal at the GPU's own hit point, FP64, with its one-ulp sensitivity.
    p = h["position"][both].astype(np.float64)
    n64 = bind.normal(cfg, p)
    n_err = np.linalg.norm(h["normal"][both] - n64, axis=-1)
    sens = np.zeros(len(p))
    for ax in range(3):
        q = p.copy()
        q[:, ax] += np.abs(p).max(axis=1) * 2.0 ** -23
        sens = np.maximum(sens, np.linalg.norm(bind.normal(cfg, q) - n64, axis=1))
    ok = n_err <= np.maximum(1e-4, 4.0 * sens)
    print(case, "normal: frac<=1e-4 %.5f within conditioning %.5f median %.2e" % (
        np.mean(n_err <= 1e-4), np.mean(ok), np.median(n_err)))
    assert np.mean(ok) >= 0.985


def test_field_samples_match_fp64_within_mandel_parity_tolerance():
    rng = np.random.default_rng(0)
    pts = rng.uniform(-1.5, 1.5, (8192, 3))
    for cfg in (scenes.mandelbulb_scene(8, 8), scenes.typed_program_scene(8, 8)):
        got = api.sample_field(cfg, pts)
        want = bind.map_distance(cfg, pts.astype(np.float32).astype(np.float64), 64)
        err = np.abs(got[:, 0] - want) / (1.0 + np.abs(want))
        assert np.all(np.isfinite(got[:, 0]))
        assert err.max() <= 5e-5, err.max()


def test_leaf_orbit_stats_match_oracle():
    cfg = scenes.mandelbulb_scene(8, 8)
    pts = np.random.default_rng(1).uniform(-1.2, 1.2, (512, 3)).astype(np.float32).astype(np.float64)
    got = api.sample_field(cfg, pts)
    it = np.array([bind.leaf_orbit(ffi.OP_MANDELBULB, [8, 12, 2, 0], p)["iterations"] for p in pts])
    assert np.mean(got[:, 3] == it) >= 0.99


def test_program_gradient_matches_fp64():
    rng = np.random.default_rng(2)
    pts = rng.uniform(-1.3, 1.3, (4096, 3)).astype(np.float32).astype(np.float64)
    for cfg in (scenes.mandelbulb_scene(8, 8), scenes.typed_program_scene(8, 8)):
        got = api.program_surface(cfg, pts)
        want = bind.program_surface(cfg, pts, 64)
        gn = np.linalg.norm(want[:, 1:], axis=1).clip(1e-9)
        rel = np.linalg.norm(got[:, 1:] - want[:, 1:], axis=1) / gn
        # one-ulp sensitivity of the FP64 gradient (fold planes, CSG seams and
        # non-escaping orbits make it large at some points)
        sens = np.zeros(len(pts))
        for ax in range(3):
            q = pts.copy()
            q[:, ax] += np.abs(pts).max(axis=1) * 2.0 ** -23
            sens = np.maximum(sens, np.linalg.norm(bind.program_surface(cfg, q, 64)[:, 1:] - want[:, 1:], axis=1) / gn)
        ok = rel <= np.maximum(1e-4, 4.0 * sens)
        print("gradient rel err median %.2e p99 %.2e within conditioning %.4f" % (
            np.median(rel), np.percentile(rel, 99), np.mean(ok)))
        assert np.mean(ok) >= 0.985 and np.median(rel) <= 1e-5


@pytest.mark.parametrize("case", ["mandelbulb_p8", "typed_program", "exact_csg"])
def test_image_parity_against_fp64_and_fp32_twin(case):
    cfg = CASES[case]()
    cfg.samples = 4
    img = api.render(cfg)
    ref = bind.render(cfg, 64)
    twin = bind.render(cfg, 32)
    rmse = float(np.sqrt(np.mean((img[..., :3] - ref[..., :3]) ** 2)))
    rmse_twin = float(np.sqrt(np.mean((twin[..., :3] - ref[..., :3]) ** 2)))
    mae = float(np.mean(np.abs(img[..., :3] - ref[..., :3])))
    print(case, "rmse gpu %.3e twin %.3e mae %.3e" % (rmse, rmse_twin, mae))
    assert np.all(np.isfinite(img))
    assert rmse <= 1.25 * rmse_twin + 2e-3


def _plan_render(cfg, kernel):
    with api.Plan(cfg) as plan:
        plan.tune(kernel=kernel)
        buf = api.DeviceBuffer(cfg.width * cfg.height * 4)
        plan.render_into(buf.address)
        return buf.numpy().reshape(cfg.height, cfg.width, 4)


def test_accumulation_policy_does_not_change_the_image():
    """Batch vs chunked vs per-sample launches of the one-path-per-lane kernel
    (metal_bridge.mm:4531-4564; the reference's capability test requires
    batch == chunked, run_capability_tests.sh:22-28) and the persistent
    kernel's single launch per frame give one image."""
    base = scenes.mandelbulb_scene(64, 48, 6)
    imgs = []
    for mode in (1, 2, 3):
        c = scenes.clone(base)
        c.sdf_accumulation_mode = mode
        c.sdf_chunk_samples = 4
        imgs.append(_plan_render(c, 0))
        imgs.append(_plan_render(c, 1))
    imgs.append(api.render(base))
    for i, im in enumerate(imgs[1:]):
        assert np.array_equal(imgs[0], im), i


def test_spp_changes_the_image():
    """run_capability_tests.sh: 1 spp != 16 spp."""
    a = api.render(scenes.mandelbulb_scene(32, 32, 1))
    b = api.render(scenes.mandelbulb_scene(32, 32, 16))
    assert not np.array_equal(a, b)


def test_profile_counters_match_oracle_work():
    cfg = scenes.mandelbulb_scene(64, 64, 2)
    res = api.render_full(cfg, profile=True)
    bind.render(cfg, 64)
    k = bind.counters()
    p = res.profile
    assert p["pixels"] == 64 * 64
    for key, okey in (("distance_evals", "distance_evals"), ("primary_steps", "primary_steps"),
                      ("shadow_steps", "shadow_steps"), ("normal_evals", "normal_evals")):
        assert abs(p[key] - k[okey]) <= 0.01 * k[okey], (key, p[key], k[okey])


def test_png_output_and_blank_guard(tmp_path):
    out = tmp_path / "m.png"
    res = api.render_full(scenes.mandelbulb_scene(64, 64, 1), output_png=str(out))
    data = out.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n" and len(data) > 100
    assert res.elapsed_ms > 0
    # Everything misses and the environment is off: a single-colour image.
    blank = scenes.mandelbulb_scene(16, 16, 1)
    blank.camera_position[2] = -50.0
    blank.camera_fov = 1.0
    blank.camera_yaw_pitch[0] = 3.0
    with pytest.raises(ffi.FptError, match="blank or single-colour"):
        api.render_full(blank, output_png=str(tmp_path / "b.png"))


@pytest.mark.parametrize("power,iterations,bailout", [(8.0, 12, 2.0), (3.0, 9, 2.0), (7.5, 12, 2.0)])
def test_single_mandelbulb_leaf_matches_the_generic_program_march(power, iterations, bailout):
    """A program that is one Mandelbulb leaf runs the persistent kernel's
    interleaved-orbit instance (one orbit iteration per trip, finalisation
    batched over finished lanes); prefixing a zero translation makes the
    same field a two-op program, which runs the generic program march.  The
    images and the work counters must be identical: integer power 8 and 3
    take the polynomial orbit step, 7.5 the trigonometric one."""
    cfg = scenes.mandelbulb_scene(72, 56, 3, bounces=2, power=power, iterations=iterations, bailout=bailout)
    generic = scenes.clone(cfg)
    scenes.set_program(generic, [(ffi.OP_TRANSLATE, 0, [0.0, 0.0, 0.0, 0.0]),
                                 (ffi.OP_MANDELBULB, 0, [power, float(iterations), bailout, 0.0])])
    assert np.array_equal(api.render(cfg), api.render(generic))
    a = api.render_full(cfg, profile=True)
    b = api.render_full(generic, profile=True)
    assert np.array_equal(a.linear, b.linear)
    assert np.isfinite(a.linear).all() and a.linear[..., :3].max() > 0
    for key in ("distance_evals", "march_orbit_iterations", "primary_steps", "shadow_steps", "normal_evals"):
        assert a.profile[key] == b.profile[key], key


@pytest.mark.parametrize("scene,spp,accumulation", [("bulb", 35, ffi.SDF_ACCUMULATION_BATCH),
                                                    ("typed", 33, ffi.SDF_ACCUMULATION_BATCH),
                                                    ("typed", 9, ffi.SDF_ACCUMULATION_BATCH),
                                                    ("bulb", 37, ffi.SDF_ACCUMULATION_CHUNKED)])
def test_path_stash_and_sample_runs_do_not_change_images(scene, spp, accumulation):
    """The path stash's knobs (refill threshold, event batch counted over a
    wave's lanes and parked paths) and the sample runs (a pixel's samples
    split into runs rendered by different lanes, the last run folding the
    per-sample colours in order) change only which lane runs which path:
    every setting renders the same frame as the one-path kernel, with
    chunked launches that resume the running mean from the buffer too, and
    out-of-range settings are refused."""
    if scene == "bulb":
        cfg = scenes.mandelbulb_scene(40, 32, spp, bounces=2)
    else:
        cfg = scenes.typed_program_scene(40, 24, spp)
    cfg.sdf_accumulation_mode = accumulation
    cfg.sdf_chunk_samples = 33  # chunked: a launch of 33 samples (runs), then one of 4 (one run)
    buf = api.DeviceBuffer(cfg.width * cfg.height * 4)
    with api.Plan(cfg) as plan:
        plan.tune(kernel=0)
        plan.render_into(buf.address)
        ref = buf.numpy().copy()
        for refill, runs, eb, wg in ((1, 1, 1, 0), (64, 64, 64, 0), (4, 4, 0, 3), (17, 3, 33, 1), (4, 2, 0, 0),
                                     (4, 8, 0, 0), (8, 8, 128, 0), (4, 5, 97, 2)):
            plan.tune(kernel=1, workgroups=wg, event_batch=eb)
            plan.tune_stash(refill, runs)
            plan.render_into(buf.address)
            assert np.array_equal(buf.numpy(), ref), (refill, runs, eb, wg)
        with pytest.raises(ffi.FptError, match="launch shape"):
            plan.tune(kernel=1, event_batch=129)
        for bad in (0, 65):
            with pytest.raises(ffi.FptError, match="refill"):
                plan.tune_stash(bad, 4)
            with pytest.raises(ffi.FptError, match="sample run"):
                plan.tune_stash(4, bad)


@pytest.mark.parametrize("spp,candidates", [(20, 4), (16, 8)])
def test_plan_autotune_picks_a_valid_shape_and_keeps_the_image(spp, candidates):
    """The single-scene autotuner (parked paths: 1-2 workgroups per compute
    unit x two tile orders, the event batch left at the launch's default,
    timed on the frame's first 16 samples; where those are the whole frame,
    also one slot per pixel against the sample runs) stores a shape that
    renders the same frame."""
    cfg = scenes.mandelbulb_scene(48, 40, spp, bounces=2)
    buf = api.DeviceBuffer(cfg.width * cfg.height * 4)
    st = api.Stream()
    with api.Plan(cfg) as plan:
        plan.render_into(buf.address)
        before = buf.numpy().copy()
        r = plan.autotune(buf.address, st.address, rounds=1)
        assert r.candidates == candidates and r.best_ms > 0 and r.event_batch == 0 and r.workgroups > 0
        assert r.tile_order in (0, 1) and r.sample_groups in ((8,) if candidates == 4 else (1, 8))
        plan.render_into(buf.address)
        assert np.array_equal(buf.numpy(), before)
