"""Voxel part of the oracle (no GPU): the reference's DDA known answers,
FP16 conversion, staging-order compaction and shell coverage invariants."""
import numpy as np
import pytest

from fractalpathtracer_amd import ffi, scenes
from oracle import bind


def test_dda_contract_known_answers_hold():
    """Shaders.metal:5156-5386 checks 0-9 (cell ownership, tied advance,
    next-float entry, slab edges) and 27-28 (precision offsets)."""
    assert bind.voxel_dda_contract() == 0


def test_f32_to_f16_matches_numpy_round_to_nearest_even():
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-5, 500), rng.normal(0, 3e4, 500),
                           [0.0, -0.0, 65504.0, 65519.0, 65520.0, 1e20, -1e20, 2 ** -24, 2 ** -25,
                            3 * 2 ** -26, 5.960464477539063e-08]]).astype(np.float32)
    got = np.array([bind.f32_to_f16_bits(float(v)) for v in vals], np.uint16)
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


def _small(storage, coverage=0, res=48):
    c = scenes.voxel_mandelbox_scene(32, 18, 1, resolution=res)
    c.voxel_storage = storage
    c.voxel_coverage_mode = coverage
    return c


def _dense_occupancy(cfg):
    c = scenes.clone(cfg)
    c.voxel_storage = ffi.VOXEL_STORAGE_DENSE
    v = bind.voxel_build(c)
    res = c.voxel_resolution
    return ((v["records"]["packed_color"] >> 31) & 1).astype(bool).reshape(res, res, res), v


@pytest.mark.parametrize("storage,B", [(ffi.VOXEL_STORAGE_SPARSE_BRICKS, 4), (ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 8)])
@pytest.mark.parametrize("coverage", [0, 3])
def test_sparse_pages_follow_staging_scan_order(storage, B, coverage):
    """Pages are numbered 1.. in linear brick order bx + G*(by + G*bz)
    (metal_bridge.mm:1744-1770) and hold exactly the dense grid's cells."""
    cfg = _small(storage, coverage, res=44)  # not a multiple of the brick size
    occ, dense = _dense_occupancy(cfg)
    v = bind.voxel_build(cfg)
    res, G = cfg.voxel_resolution, v["bricks_per_axis"]
    assert G == -(-res // B)
    pad = np.zeros((G * B,) * 3, bool)
    pad[:res, :res, :res] = occ
    brick_occ = pad.reshape(G, B, G, B, G, B).any(axis=(1, 3, 5)).reshape(-1)
    want = np.zeros(G ** 3, np.uint32)
    want[brick_occ] = np.arange(1, brick_occ.sum() + 1)
    assert np.array_equal(v["page_table"], want)
    assert v["active_cells"] == occ.sum() == dense["active_cells"]
    # cell payloads: occupied flag of every stored cell equals the dense grid's
    z, y, x = np.nonzero(pad.reshape(G, B, G, B, G, B).transpose(0, 2, 4, 1, 3, 5).any(axis=(3, 4, 5)))
    for bz, by, bx in list(zip(z, y, x))[:50]:
        page = v["page_table"][bx + G * (by + G * bz)]
        block = pad[bz * B:(bz + 1) * B, by * B:(by + 1) * B, bx * B:(bx + 1) * B]  # (lz, ly, lx)
        if B == 8:
            stored = v["material"][(page - 1) * 512:page * 512].reshape(8, 8, 8) != 0
        else:
            stored = (v["records"]["packed_color"][(page - 1) * 64:page * 64].reshape(4, 4, 4) >> 31) == 1
        assert np.array_equal(stored, block)


def test_shell_drops_exactly_the_fully_enclosed_cells():
    legacy, _ = _dense_occupancy(_small(ffi.VOXEL_STORAGE_DENSE, 0, res=40))
    shell, _ = _dense_occupancy(_small(ffi.VOXEL_STORAGE_DENSE, 3, res=40))
    res = legacy.shape[0]
    pad = np.pad(legacy, 1, constant_values=False)
    enclosed = np.ones_like(legacy)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                enclosed &= pad[1 + dz:1 + dz + res, 1 + dy:1 + dy + res, 1 + dx:1 + dx + res]
    assert np.array_equal(shell, legacy & ~enclosed)
    assert shell.sum() < 0.2 * legacy.sum()


def test_shell_and_legacy_grids_give_identical_first_hits_from_outside():
    rng = np.random.default_rng(5)
    n = 400
    origins = rng.uniform(-1, 1, (n, 3))
    origins = origins / np.linalg.norm(origins, axis=1, keepdims=True) * 12.0
    targets = rng.uniform(-5, 5, (n, 3))
    dirs = targets - origins
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays = np.concatenate([origins, dirs], axis=1).astype(np.float32)
    a = bind.voxel_trace_cells(_small(ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 0, 40), rays, cap=8)[2]
    b = bind.voxel_trace_cells(_small(ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 3, 40), rays, cap=8)[2]
    assert np.array_equal(a[:, :6], b[:, :6])
    assert a[:, 0].sum() > 100


def test_dda_visits_are_face_or_tie_connected_and_monotone():
    cfg = _small(ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 3, 40)
    rng = np.random.default_rng(6)
    o = rng.uniform(-8, 8, (200, 3))
    d = rng.normal(0, 1, (200, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    cells, counts, hits = bind.voxel_trace_cells(cfg, np.concatenate([o, d], 1), cap=512)
    for i in range(200):
        k = min(int(counts[i]), 512)
        if k < 2:
            continue
        steps = np.diff(cells[i, :k], axis=0)
        assert np.all(np.abs(steps) <= 1) and np.all(np.abs(steps).sum(axis=1) >= 1)
        # every axis moves in one direction only
        for a in range(3):
            s = steps[:, a]
            assert not (np.any(s > 0) and np.any(s < 0))


@pytest.mark.parametrize("storage,B", [(ffi.VOXEL_STORAGE_SPARSE_BRICKS, 4), (ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 8),
                                       (ffi.VOXEL_STORAGE_DENSE, 0)])
def test_memory_work_counts_follow_the_visit_lists(storage, B):
    """The counting pass's definition of the DDA's memory work, recomputed in
    numpy from each ray's visit list and the built page table: every visited
    cell counts once; a page-table entry is read whenever the visited brick
    differs from the previous one; a cell payload is read for every visited
    cell in an occupied page (every cell of a dense grid)."""
    cfg = _small(storage, 3 if storage != ffi.VOXEL_STORAGE_DENSE else 0, 40)
    rng = np.random.default_rng(9)
    o = rng.uniform(-9, 9, (300, 3))
    d = rng.normal(0, 1, (300, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    cells, counts, _ = bind.voxel_trace_cells(cfg, np.concatenate([o, d], 1), cap=512)
    got = bind.voxel_counts()
    v = bind.voxel_build(cfg)
    visited = pages = payload = 0
    for i in range(len(o)):
        k = int(counts[i])
        assert k <= 512
        visited += k
        if k == 0:
            continue
        if storage == ffi.VOXEL_STORAGE_DENSE:
            payload += k
            continue
        G = v["bricks_per_axis"]
        b = cells[i, :k] // B
        bid = b[:, 0] + G * (b[:, 1] + G * b[:, 2])
        pages += 1 + int(np.count_nonzero(np.diff(bid)))
        payload += int(np.count_nonzero(v["page_table"][bid]))
    assert visited > 100 and (storage == ffi.VOXEL_STORAGE_DENSE or 0 < payload < visited)
    assert got == {"visited_cells": visited, "page_reads": pages, "cell_reads": payload}
