"""Learned step proposer: parameter layout, the reference network's pins,
and (GPU) the MFMA network against the reference, the training sampler, and
learned-step renders against exact-step renders and the FP64 oracle."""
import os
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from fractalpathtracer_amd import ffi, proposer
from oracle import proposer_ref

ROOT = Path(__file__).resolve().parents[1]


# --------------------------------------------------------------------------
# CPU: layout and reference pins.
# --------------------------------------------------------------------------

def test_param_count_matches_header_and_binding():
    text = (ROOT / "include" / "fpt_abi.h").read_text()
    m = re.search(r"#define FPT_PROPOSER_PARAM_COUNT (\d+)", text)
    assert m and int(m.group(1)) == proposer.PARAM_COUNT == ffi.PROPOSER_PARAM_COUNT
    # 128*64 + 64 + 3*(64*64 + 64) + 64 + 1
    assert proposer.PARAM_COUNT == 20801
    lay = proposer.layout()
    assert lay["b5"][0] == proposer.PARAM_COUNT - 1 and lay["W2"][0] == 64 * 128 + 64


def test_split_join_and_weight_file_round_trip(tmp_path):
    p = proposer.init_params(7)
    assert np.array_equal(proposer.join(proposer.split(p)), p)
    proposer.save(tmp_path / "w.npz", p, {"note": "x"})
    q, meta = proposer.load(tmp_path / "w.npz")
    assert np.array_equal(p, q) and meta == {"note": "x"}
    with pytest.raises(ValueError):
        proposer.split(p[:-1])


def test_bf16_round_known_bit_patterns():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 has 7 fraction bits
    x = np.array([one, one + 2.0 ** -8, one + 3 * 2.0 ** -8, one + 2.0 ** -8 + 2.0 ** -20, -2.5,
                  3.0e38, 1.0e-40], np.float32)
    want = np.array([one, one, one + 2 * ulp, one + ulp, -2.5, 2.9980e38, 1.0e-40], np.float64)
    got = proposer_ref.bf16_round(x)
    # ties go to the even mantissa; just above the tie rounds up
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == one + 2 * ulp and got[3] == one + ulp
    assert got[4] == -2.5
    assert np.allclose(got[5], want[5], rtol=2 ** -8)
    assert got.view(np.uint32)[6] & 0xFFFF == 0  # low half cleared for subnormals too


def test_features_closed_forms():
    h = 1.5
    f0 = proposer_ref.features(np.zeros((1, 3)), h)[0]
    for a in range(3):
        for k in range(16):
            assert f0[32 * a + 2 * k] == 0.0 and f0[32 * a + 2 * k + 1] == 1.0
    assert np.all(f0[96:] == 0.0)
    # v = 1: sin(pi 2^m) = 0, cos(pi) = -1, cos(pi 2^m) = 1 for m >= 1;
    # odd k: frequency 2^m sqrt2 with sin/cos of pi 2^m sqrt2.
    f1 = proposer_ref.features(np.array([[h, 0.0, 0.0]]), h)[0]
    for k in range(16):
        w = np.pi * 2.0 ** (k / 2.0)
        assert abs(f1[2 * k] - np.sin(w)) < 1e-12 and abs(f1[2 * k + 1] - np.cos(w)) < 1e-12
        if k % 2 == 0:
            assert abs(f1[2 * k]) < 1e-9
            assert f1[2 * k + 1] == pytest.approx(-1.0 if k == 0 else 1.0)
    assert f1[96] == 1.0 and f1[97] == 0.0 and f1[98] == 0.0


def test_reference_forward_matches_training_definition():
    """numpy reference == PyTorch training network (quantised forward), run
    in a child process so PyTorch's HIP runtime stays out of this one."""
    params = proposer.init_params(3)
    rng = np.random.default_rng(1)
    pts = rng.uniform(-1.7, 1.7, (257, 3))
    ref = proposer_ref.forward(params, pts, 1.5)
    script = (
        "import sys, numpy as np, torch\n"
        "from fractalpathtracer_amd import proposer as P, proposer_train as T\n"
        "p = {k: torch.tensor(v, dtype=torch.float64) for k, v in P.split(np.load(sys.argv[1])).items()}\n"
        "v = torch.tensor(np.load(sys.argv[2]) / 1.5, dtype=torch.float64)\n"
        "def q(t):\n"
        "    return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)\n"
        "T._bf16 = q\n"
        "np.save(sys.argv[3], T.forward(p, v, True).detach().numpy())\n")
    d = Path(os.environ.get("TMPDIR", "/tmp"))
    tag = f"fpt_prop_{os.getpid()}"
    files = [d / f"{tag}_{n}.npy" for n in ("p", "x", "o")]
    np.save(files[0], params)
    np.save(files[1], pts)
    env = dict(os.environ, PYTHONPATH=str(ROOT), CUDA_VISIBLE_DEVICES="", HIP_VISIBLE_DEVICES="")
    try:
        r = subprocess.run([sys.executable, "-c", script, *map(str, files)], env=env, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        out = np.load(files[2])
    finally:
        for f in files:
            f.unlink(missing_ok=True)
    # Both round features/activations through BF16 from (nearly) the same
    # float64 values; rare ties may round apart, so compare loosely per point
    # and tightly on the median.
    err = np.abs(out - ref)
    assert np.median(err) < 1e-6 and err.max() < 5e-2 * (1 + np.abs(ref).max())


# --------------------------------------------------------------------------
# GPU.
# --------------------------------------------------------------------------

def _plan_image(cfg, params=None, **settings):
    from fractalpathtracer_amd import api
    with api.Plan(cfg) as plan:
        if params is not None:
            plan.set_proposer(params, **settings)
        buf = api.DeviceBuffer(cfg.width * cfg.height * 4)
        plan.render_into(buf.address)
        img = buf.numpy().reshape(cfg.height, cfg.width, 4)
        prof, learned = plan.profile(buf)
        buf.close()
    return img, prof, learned


def _constant_proposer(value):
    parts = {k: np.zeros(s) for k, (_, s) in proposer.layout().items()}
    parts["b5"] = np.array([value])
    return proposer.join(parts)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 63, 64, 1000, 4099])
def test_gpu_proposer_matches_reference_network(n):
    from fractalpathtracer_amd import api
    params = proposer.init_params(11)
    rng = np.random.default_rng(n)
    pts = rng.uniform(-1.8, 1.8, (n, 3)).astype(np.float32)
    got = api.proposer_eval(params, pts, 1.5)
    ref = proposer_ref.forward(params, pts.astype(np.float64), 1.5)
    err = np.abs(got - ref)
    scale = 1.0 + np.abs(ref)
    print(f"n={n} max err {err.max():.3e} median {np.median(err):.3e} |ref| max {np.abs(ref).max():.3f}")
    # GPU: hardware sine, exp and reciprocal, FP32 MFMA sums in hardware order;
    # a BF16 input that rounds apart moves one product by 2^-8 relative.
    assert np.median(err / scale) < 2e-3
    assert np.max(err / scale) < 3e-2


@pytest.mark.gpu
def test_gpu_proposer_rejects_bad_parameters():
    from fractalpathtracer_amd import api
    with pytest.raises(ffi.FptError, match="FPT_PROPOSER_PARAM_COUNT"):
        api.proposer_eval(np.zeros(10, np.float32), np.zeros((4, 3)))
    bad = proposer.init_params(0)
    bad[5] = np.nan
    with pytest.raises(ffi.FptError, match="finite"):
        api.proposer_eval(bad, np.zeros((4, 3)))


@pytest.mark.gpu
def test_gpu_training_samples():
    from fractalpathtracer_amd import api, scenes
    from oracle import bind
    cfg = scenes.learned_step_scene(64, 64, 1)
    n = 1 << 14
    pts, fails = api.proposer_samples(cfg, n, seed=5, box_half=1.5, band=0.05)
    again, _ = api.proposer_samples(cfg, n, seed=5, box_half=1.5, band=0.05)
    other, _ = api.proposer_samples(cfg, n, seed=6, box_half=1.5, band=0.05)
    assert fails == 0
    assert np.array_equal(pts, again) and not np.array_equal(pts[:, :3], other[:, :3])
    assert np.all(np.abs(pts[:, :3]) <= 1.5)
    uni, near = pts[: n // 2], pts[n // 2:]
    # uniform half: mean 0, variance h^2/3 per axis
    assert np.all(np.abs(uni[:, :3].mean(0)) < 0.05)
    assert np.allclose(uni[:, :3].var(0), 0.75, rtol=0.05)
    assert np.all(np.abs(near[:, 3]) < 0.05)
    # labels are the scene's distance (FP32 GPU vs FP64 oracle field tolerance)
    sel = np.r_[0:256, n // 2:n // 2 + 256]
    ref = bind.map_distance(cfg, pts[sel, :3].astype(np.float64), 64)
    assert np.all(np.abs(pts[sel, 3] - ref) <= 5e-5 * (1 + np.abs(ref)))


@pytest.mark.gpu
def test_gpu_proposer_below_switch_gives_the_exact_image():
    """A proposer that always answers below the switch distance hands every
    march straight back to the exact distance: the image is bit-identical."""
    from fractalpathtracer_amd import scenes
    cfg = scenes.learned_step_scene(96, 72, 3)
    exact, prof_e, _ = _plan_image(cfg)
    img, prof_l, learned = _plan_image(cfg, _constant_proposer(0.01))
    assert np.array_equal(img, exact)
    assert learned["proposals"] > 0 and learned["far_steps"] == 0 and learned["backtracks"] == 0
    assert prof_l["distance_evals"] == prof_e["distance_evals"]


@pytest.mark.gpu
def test_gpu_constant_far_proposer_terminates():
    """A proposal that never drops to the switch distance keeps every march
    on far steps inside the box; the far-step budget then hands it to the
    exact distance, so every pixel still finishes with a finite value."""
    from fractalpathtracer_amd import scenes
    cfg = scenes.learned_step_scene(96, 72, 2)
    img, prof, learned = _plan_image(cfg, _constant_proposer(0.03))
    assert learned["far_steps"] > 0 and learned["backtracks"] == 0
    assert prof["pixels"] == cfg.width * cfg.height
    assert np.all(np.isfinite(img)) and np.all(img[..., 3] == 1.0)


@pytest.mark.gpu
def test_gpu_trained_proposer_render_parity():
    """Learned stepping moves hit points within the exact march's convergence
    threshold, which decorrelates its paths from the oracle's same-RNG paths
    but must not bias the image: both methods are scored against an
    independent FP64 estimate (oracle samples spp .. 17*spp - 1) and must have
    equal error (within 10%) and the same mean."""
    from fractalpathtracer_amd import scenes
    from oracle import bind
    path = proposer.asset_path("mandelbulb_p8")
    if not path.exists():
        pytest.skip("trained Mandelbulb proposer weights are not present")
    params, meta = proposer.load(path)
    cfg = scenes.learned_step_scene(80, 60, 4)
    exact, _, _ = _plan_image(cfg)
    img, _, learned = _plan_image(cfg, params, box_half=meta.get("box_half", 1.5))
    conv = bind.render_samples(cfg, (4, 68))
    rm = lambda a: float(np.sqrt(np.mean((a[..., :3] - conv[..., :3]) ** 2)))  # noqa: E731
    print("rmse vs converged: exact %.4e learned %.4e; means %.5f %.5f %.5f" % (
        rm(exact), rm(img), exact[..., :3].mean(), img[..., :3].mean(), conv[..., :3].mean()), learned)
    assert learned["far_steps"] > 0 and learned["proposals"] >= learned["far_steps"]
    assert rm(img) <= 1.10 * rm(exact)
    assert abs(img[..., :3].mean() - exact[..., :3].mean()) < 0.02 * conv[..., :3].mean()
