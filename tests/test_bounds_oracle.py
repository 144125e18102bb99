"""Pins of the oracle's program bounds and cell classification
(Shaders.metal:1065-1273, 4196-4244): the reference's own known answers from
its contract test (Shaders.metal:5215-5298 on the programs of
metal_bridge.mm:3748-3805), the sphere's closed-form interval, and
containment of brute-force distance samples over random boxes and
programs."""
import itertools

import numpy as np
import pytest

from fractalpathtracer_amd import ffi, scenes
from oracle import bind

S, B, P = ffi.OP_SPHERE, ffi.OP_BOX, ffi.OP_PLANE
OUTSIDE, INSIDE, SURFACE, UNKNOWN = 0, 1, 2, 3


def contract_programs():
    """The nine typed programs of the reference's bounds contract."""
    progs = [
        [(S, 0, [1.0])],
        [(B, 0, [0.8, 0.6, 0.4])],
        [(P, 0, [1.0, 2.0, -0.5, 0.2])],
        [(ffi.OP_REPEAT, 0, [4.0, 4.0, 4.0]), (S, 0, [1.0])],
        [(ffi.OP_ABS, 0, []), (ffi.OP_ROTATE_X, 0, [0.31]), (ffi.OP_ROTATE_Y, 0, [-0.47]),
         (ffi.OP_ROTATE_Z, 0, [0.19]), (ffi.OP_SCALE, 0, [1.25]), (ffi.OP_TRANSLATE, 0, [0.2, -0.1, 0.3]),
         (B, 0, [1.0, 0.8, 0.6]), (S, 2, [0.35])],
        [(S, 0, [1.0]), (ffi.OP_TRANSLATE, 0, [0.5, 0, 0]), (B, 1, [0.7, 0.7, 0.7])],
        [(S, 0, [0.7]), (ffi.OP_TRANSLATE, 0, [0.9, 0, 0]), (S, 0, [0.7])],
        [(ffi.OP_SORT_DESC, 0, []), (S, 0, [1.0])],
        [(ffi.OP_SORT_DESC, 0, []), (B, 0, [0.9, 0.55, 0.3])],
    ]
    out = []
    for ops in progs:
        c = scenes.default_config()
        scenes.set_program(c, ops)
        c.voxel_surface_band = 0.5
        out.append(c)
    return out


def contains_grid(cfg, center, half, n, tol, precision):
    b = bind.program_bounds(cfg, center, half, precision)
    if not b["certified"]:
        return False
    u = np.linspace(-1, 1, n)
    pts = np.array([np.asarray(center) + np.array(t) * np.asarray(half) for t in itertools.product(u, u, u)])
    d = bind.map_distance(cfg, pts, precision)
    return bool(np.all(np.isfinite(d)) and np.all(d >= b["minimum"] - tol) and np.all(d <= b["maximum"] + tol))


@pytest.mark.parametrize("precision", [32, 64])
def test_reference_bounds_contract(precision):
    """Bits 10-21 and 29 of the reference's voxel/bounds contract."""
    pc = contract_programs()
    q = [0.25] * 3
    sb = bind.program_bounds(pc[0], [0, 0, 0], q, precision)
    inside = bind.classify_cell(pc[0], [0, 0, 0], q, precision)
    outside = bind.classify_cell(pc[0], [2, 0, 0], q, precision)
    surface = bind.classify_cell(pc[0], [1, 0, 0], [0.1] * 3, precision)
    # bit 11
    assert sb["certified"] and sb["lipschitz"] == 1.0 and inside["certified"]
    assert inside["cell_class"] == INSIDE and outside["cell_class"] == OUTSIDE and surface["cell_class"] == SURFACE
    assert contains_grid(pc[0], [0, 0, 0], q, 3, 2e-5, precision)
    # bit 20
    assert surface["lipschitz"] and surface["interval"] and not inside["lipschitz"] and not inside["interval"]
    # bit 21: legacy misses a cell the certified tests keep
    miss = bind.classify_cell(pc[0], [0.75] * 3, q, precision)
    assert not miss["legacy"] and miss["lipschitz"] and miss["interval"] and miss["cell_class"] == SURFACE
    # bit 19
    assert not outside["legacy"] and surface["legacy"]
    # bits 12-17
    assert contains_grid(pc[1], [0.35, -0.15, 0.2], [0.4, 0.3, 0.2], 3, 2e-5, precision)
    assert bind.program_bounds(pc[2], [0.2, -0.1, 0.3], [0.4, 0.2, 0.5], precision)["certified"]
    assert contains_grid(pc[2], [0.2, -0.1, 0.3], [0.4, 0.2, 0.5], 3, 2e-5, precision)
    assert bind.program_bounds(pc[3], [0, 0, 0], [0.3] * 3, precision)["certified"]
    assert not bind.program_bounds(pc[3], [2, 0, 0], [0.1] * 3, precision)["certified"]   # straddles a seam
    assert contains_grid(pc[3], [0, 0, 0], [0.3] * 3, 3, 2e-5, precision)
    assert contains_grid(pc[4], [0.4, -0.2, 0.1], [0.6, 0.5, 0.4], 3, 2e-5, precision)
    assert contains_grid(pc[5], [0.2, 0, 0], [0.8, 0.5, 0.4], 3, 2e-5, precision)
    assert contains_grid(pc[6], [0.4, 0, 0], [1.1, 0.5, 0.4], 3, 2e-5, precision)
    # bit 18
    sorted_surface = bind.classify_cell(pc[7], [1, 0, 0], [0.1] * 3, precision)
    assert bind.program_bounds(pc[7], [0, 0, 0], q, precision)["certified"]
    assert sorted_surface["cell_class"] == SURFACE and sorted_surface["certified"]
    assert sorted_surface["legacy"] and sorted_surface["lipschitz"] and sorted_surface["interval"]
    assert contains_grid(pc[7], [0, 0, 0], q, 3, 2e-5, precision)
    # bit 29
    assert contains_grid(pc[8], [0.2, -0.35, 0.1], [0.7, 0.45, 0.55], 3, 2e-5, precision)


def test_sphere_interval_closed_form():
    c = contract_programs()[0]
    for center, half in [([0.3, -0.2, 0.1], [0.2, 0.1, 0.3]), ([2, 1, 0], [0.5, 0.5, 0.5]), ([0, 0, 0], [2, 2, 2])]:
        b = bind.program_bounds(c, center, half)
        ac = np.abs(center)
        lo = np.linalg.norm(np.maximum(ac - half, 0)) - 1.0
        hi = np.linalg.norm(ac + np.asarray(half)) - 1.0
        # widened outward by the allowance (>= 2e-5 (1 + |d|)) and one ulp
        assert lo - 1e-3 < b["minimum"] < lo and hi < b["maximum"] < hi + 1e-3


def test_uncertified_programs():
    c = scenes.default_config()
    scenes.set_program(c, [(ffi.OP_MANDELBULB, 0, [8, 12, 2])])
    assert not bind.program_bounds(c, [0, 0, 0], [0.1] * 3)["certified"]
    k = bind.classify_cell(c, [0, 0, 0], [0.1] * 3)
    assert k["cell_class"] == UNKNOWN and k["legacy"] == k["lipschitz"] == k["interval"]
    scenes.set_program(c, [(P, 0, [0, 0, 0, 1])])   # zero normal: no bounds
    assert not bind.program_bounds(c, [0, 0, 0], [0.1] * 3)["certified"]


def test_bounds_contain_dense_samples_on_random_boxes():
    """Brute force: min/max of 7^3 samples of the program distance stay
    inside the certified interval, for random programs of the bounded ops."""
    rng = np.random.default_rng(7)
    ops_pool = [ffi.OP_ABS, ffi.OP_TRANSLATE, ffi.OP_SCALE, ffi.OP_ROTATE_X, ffi.OP_ROTATE_Y, ffi.OP_ROTATE_Z,
                ffi.OP_SORT_DESC]
    checked = 0
    for trial in range(60):
        ops = []
        for _ in range(rng.integers(0, 4)):
            op = int(rng.choice(ops_pool))
            data = {ffi.OP_TRANSLATE: list(rng.uniform(-0.5, 0.5, 3)), ffi.OP_SCALE: [float(rng.uniform(0.5, 2.0))],
                    ffi.OP_ROTATE_X: [float(rng.uniform(-3, 3))], ffi.OP_ROTATE_Y: [float(rng.uniform(-3, 3))],
                    ffi.OP_ROTATE_Z: [float(rng.uniform(-3, 3))]}.get(op, [])
            ops.append((op, 0, data))
        for j in range(rng.integers(1, 4)):
            kind = int(rng.choice([S, B, P]))
            data = [float(rng.uniform(0.3, 1.0))] if kind == S else (
                list(rng.uniform(0.2, 0.9, 3)) if kind == B else list(rng.uniform(-1, 1, 3)) + [float(rng.uniform(-.5, .5))])
            ops.append((kind, int(rng.integers(0, 3)) if j else 0, data))
        c = scenes.default_config()
        scenes.set_program(c, ops)
        center = rng.uniform(-1.2, 1.2, 3)
        half = rng.uniform(0.02, 0.5, 3)
        for precision in (32, 64):
            b = bind.program_bounds(c, center, half, precision)
            if not b["certified"]:
                continue
            u = np.linspace(-1, 1, 7)
            pts = np.array([center + np.array(t) * half for t in itertools.product(u, u, u)])
            d = bind.map_distance(c, pts, 64)
            assert d.min() >= b["minimum"] and d.max() <= b["maximum"], (trial, ops)
            checked += 1
    assert checked > 60


def test_interval_and_lipschitz_coverage_keep_every_surface_cell():
    """Voxel build (oracle, dense): every cell whose 3x3x3 distance samples
    change sign holds surface and must be occupied under interval and
    Lipschitz coverage (legacy's proximity band may keep or drop such a cell
    depending on the band)."""
    c = scenes.voxel_mandelbox_scene(16, 16, 1, resolution=20)
    scenes.set_program(c, [(ffi.OP_ABS, 0, []), (ffi.OP_TRANSLATE, 0, [0.2, -0.1, 0.3]),
                           (B, 0, [1.0, 0.8, 0.6]), (S, 2, [0.9])])
    for i in range(3):
        c.voxel_bounds_min[i], c.voxel_bounds_max[i] = -2.0, 2.0
    c.voxel_storage = ffi.VOXEL_STORAGE_DENSE
    res = c.voxel_resolution
    cs = 4.0 / res
    occ = {}
    for cov in (ffi.VOXEL_COVERAGE_LEGACY, ffi.VOXEL_COVERAGE_LIPSCHITZ, ffi.VOXEL_COVERAGE_INTERVAL):
        c.voxel_coverage_mode = cov
        t = bind.voxel_build(c)
        rec = t["records"].view(np.uint32).reshape(-1, 3)
        occ[cov] = (rec[:, 0] & 0x80000000) != 0
    idx = np.arange(res)
    zz, yy, xx = np.meshgrid(idx, idx, idx, indexing="ij")
    centers = np.stack([xx, yy, zz], -1).reshape(-1, 3) * cs - 2.0 + cs / 2
    u = np.linspace(-0.5, 0.5, 3) * cs
    offs = np.array(list(itertools.product(u, u, u)))
    d = bind.map_distance(c, (centers[:, None, :] + offs[None]).reshape(-1, 3), 32).reshape(len(centers), 27)
    crossing = (d.min(1) <= 0) & (d.max(1) >= 0)
    assert crossing.sum() > 100
    assert np.all(occ[ffi.VOXEL_COVERAGE_INTERVAL][crossing]) and np.all(occ[ffi.VOXEL_COVERAGE_LIPSCHITZ][crossing])
