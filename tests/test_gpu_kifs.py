"""Mandelbulber native KIFS field on the GPU (fpt_mandelbulber_sample_field,
Shaders.metal:897-949, 1001-1008) against the FP64 oracle, with the
reference's mandel-parity tolerance (main.rs:3297-3361): distance within
5e-5 (1 + |d|) at the 99th percentile, no outlier above 2e-4 (1 + |d|) among
points whose orbit length agrees."""
import math

import numpy as np
import pytest

from fractalpathtracer_amd import api, ffi, scenes
from oracle import bind

pytestmark = pytest.mark.gpu

R2 = 1.0 / math.sqrt(2.0)


def kifs_scene(world_scale=1.0, vset=None):
    c = scenes.default_config()
    c.sdf_id = ffi.SDF_MANDELBULBER
    sv = [0.0] * 40
    sv[0], sv[1], sv[2], sv[3] = world_scale, 12, 10.0, 3.0
    sv[4], sv[5] = 7, 7
    sv[6:9] = [1.0, 1.0, 1.0]
    sv[9:12] = [0.1, -0.2, 0.05]
    sv[12:21] = [R2, -R2, 0, R2, 0, -R2, 0, R2, -R2]
    for i, v in enumerate(sv):
        c.set_values[i] = v
    for i, v in (vset or {}).items():
        c.vset_values[i] = v
    return c


@pytest.mark.parametrize("world_scale,vset", [(1.0, {}), (2.5, {120: 0.1, 123: 0.3, 125: -0.2, 126: 3.0}),
                                              (1.0, {130: 1.5, 115: 1.0, 116: 0.01, 118: 1e-3, 119: 1.0})])
def test_kifs_field_matches_oracle(world_scale, vset):
    c = kifs_scene(world_scale, vset)
    pts = (np.random.default_rng(3).uniform(-1.6, 1.6, (4096, 3)) * world_scale).astype(np.float32).astype(np.float64)
    got = api.sample_field(c, pts)
    ref = bind.kifs_sample(c, pts, 64)
    same = got[:, 3] == np.abs(ref[:, 3])
    print(f"orbit length agrees on {same.mean():.4f}")
    assert same.mean() > 0.97
    tol = 1 + np.abs(ref[same, 0])
    err = np.abs(got[same, 0] - ref[same, 0]) / tol
    assert np.quantile(err, 0.99) <= 5e-5 and err.max() <= 2e-4
    assert np.allclose(got[same, 2], ref[same, 2], rtol=1e-5)
    assert np.allclose(got[same, 1], ref[same, 1], rtol=1e-4, atol=1e-5)
