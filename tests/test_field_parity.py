"""The field-parity harness (tests/field_parity.py, a restatement of the
reference's mandel-parity command, src/main.rs:3057-3428), on the CPU.

Pins: the point hash against the published inverse of lowbias32 (Wellons,
"Prospecting for Hash Functions"; the multipliers 0x43021123 and 0x1d69e2a5
are the modular inverses of the forward ones, checked here); the fixed
sample points against the imported scene's own camera position
(mandelbulber.rs:1185-1188 writes the same FP32 world point); the
nearest-rank percentile against numpy's "nearest" method; the gates with the
oracle's FP32 twin standing in for the GPU (it qualifies) and with mistakes
planted in it (a relative distance error, a non-finite sample, a changed
orbit length), each of which the report must flag.
"""
from pathlib import Path

import numpy as np
import pytest

import field_parity as fp
from fractalpathtracer_amd import scenes

IFS20 = Path(__file__).resolve().parent / "golden" / "mandelbulber" / "ifs-20.fract"


def lowbias32_inverse(x):
    x = np.asarray(x, np.uint32).copy()
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x43021123)
    x ^= (x >> np.uint32(15)) ^ (x >> np.uint32(30))
    x *= np.uint32(0x1D69E2A5)
    return x ^ (x >> np.uint32(16))


def test_parity_hash_is_lowbias32():
    assert (0x846CA68B * 0x43021123) % 2**32 == 1 and (0x7FEB352D * 0x1D69E2A5) % 2**32 == 1
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.integers(0, 2**32, 100_000, dtype=np.uint64).astype(np.uint32),
                        np.array([0, 1, 2**31, 2**32 - 1], np.uint32)])
    h = fp.parity_hash(x)
    assert np.array_equal(lowbias32_inverse(h), x)
    assert fp.parity_hash(np.uint32(0)) == 0 and len(np.unique(h)) == len(np.unique(x))


def test_points_fixed_hashed_and_round_trip():
    cfg, _ = scenes.load_scene(IFS20)
    ws = float(cfg.set_values[0])
    cam = fp.fract_vec3(IFS20, "camera", (0, 0, -5))
    tgt = fp.fract_vec3(IFS20, "target", (0, 0, 0))
    assert tgt == [0.810487393068948, -0.405416779822347, -1.3722128396787]
    gpu, cpu = fp.parity_points(cam, tgt, ws, 4096)
    assert gpu.dtype == np.float32 and gpu.shape == (4096, 3)
    assert np.array_equal(gpu[1], np.asarray(list(cfg.camera_position)[:3], np.float32))
    assert np.array_equal(gpu[0], [0, 0, 0]) and np.array_equal(gpu[5], [ws, ws, ws])
    assert np.array_equal(gpu[6], np.float32([-ws, -ws, ws]))  # Mandelbulber (-1, 1, -1): y and z swap
    # the CPU points are the GPU's FP32 points mapped back, exactly
    assert np.array_equal((cpu[:, [0, 2, 1]] * ws).astype(np.float32), gpu)
    h = cpu[8:]
    assert h.min() >= -2 and h.max() <= 2 and np.all(np.abs(h.mean(axis=0)) < 0.1)
    with pytest.raises(ValueError):
        fp.parity_points(cam, tgt, ws, 15)


def test_percentile_nearest_rank():
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 100, 4097):
        v = rng.standard_normal(n)
        assert fp.percentile(v, 0.99) == np.percentile(v, 99, method="nearest")
    assert fp.percentile([], 0.99) == 0.0


def test_report_gates_with_twin_and_planted_mistakes():
    from oracle import bind
    cfg, _ = scenes.load_scene(IFS20)
    ws = float(cfg.set_values[0])
    gpu_pts, cpu_pts = fp.parity_points(fp.fract_vec3(IFS20, "camera", (0, 0, -5)),
                                        fp.fract_vec3(IFS20, "target", (0, 0, 0)), ws, 4096)
    ref = fp.cpu_field(cfg, cpu_pts, ws)
    twin = bind.kifs_sample(cfg, gpu_pts.astype(np.float64), 32)
    twin[:, 3] = np.abs(twin[:, 3])
    r = fp.parity_report(twin, ref, ws)
    assert r["geometry_qualified"] and r["non_finite_samples"] == 0
    assert r["p99_absolute_distance_error"] < 5e-6
    bad = twin.copy()
    bad[:, 0] *= 1 + 2e-4  # a relative error the 5e-5 bar must catch on the far samples
    rb = fp.parity_report(bad, ref, ws)
    assert not rb["geometry_qualified"] and rb["distance_failures"] > 0
    nf = twin.copy()
    nf[10, 0] = np.nan
    rn = fp.parity_report(nf, ref, ws)
    assert rn["non_finite_samples"] == 1 and not rn["qualified"]
    it = twin.copy()
    it[:3, 3] += 1
    ri = fp.parity_report(it, ref, ws)
    assert ri["iteration_mismatches"] == r["iteration_mismatches"] + 3
    assert ri["geometry_qualified"]
