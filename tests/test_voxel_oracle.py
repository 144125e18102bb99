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


# --- certified whole-brick rejection (Shaders.metal:5704-5726) --------------------

def _sphere_rejection_cfg(storage, coverage, res=32, band=0.5, fill=0, radius=1.0):
    """The reference's direct-build parity scene (metal_bridge.mm:3836-3923):
    a sphere program on [-2, 2]^3."""
    c = scenes.default_config()
    scenes.set_program(c, [(ffi.OP_SPHERE, 0, [radius])])
    c.voxel_resolution = res
    c.voxel_storage = storage
    c.voxel_coverage_mode = coverage
    c.voxel_surface_band = band
    c.voxel_fill_interior = fill
    for i in range(3):
        c.voxel_bounds_min[i], c.voxel_bounds_max[i] = -2.0, 2.0
    return c


def _cells_through_pages(v, res):
    """Occupied flag of every grid cell, looked up through the page table."""
    B, G = v["brick_size"], v["bricks_per_axis"]
    occ = np.zeros((res, res, res), bool)
    for page_index in np.nonzero(v["page_table"])[0]:
        page = int(v["page_table"][page_index]) - 1
        bx, by, bz = page_index % G, (page_index // G) % G, page_index // (G * G)
        if B == 8:
            cells = v["material"][page * 512:(page + 1) * 512] != 0
        else:
            cells = (v["records"]["packed_color"][page * 64:(page + 1) * 64] >> 31).astype(bool)
        blk = cells.reshape(B, B, B)  # (z, y, x)
        z0, y0, x0 = bz * B, by * B, bx * B
        sz, sy, sx = min(B, res - z0), min(B, res - y0), min(B, res - x0)
        occ[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = blk[:sz, :sy, :sx]
    return occ


@pytest.mark.parametrize("storage", [ffi.VOXEL_STORAGE_SPARSE_BRICKS, ffi.VOXEL_STORAGE_SPARSE_BRICKS8])
def test_direct_build_with_interval_coverage_equals_staging(storage):
    """The reference's own parity check: with interval coverage no cell of a
    brick certified off the surface is occupied, so rejection changes
    nothing but the count (metal_bridge.mm:3836-3923)."""
    cfg = _sphere_rejection_cfg(storage, ffi.VOXEL_COVERAGE_INTERVAL)
    staging = bind.voxel_build(cfg)
    cfg.voxel_build_mode, cfg.voxel_brick_rejection = 1, 1
    direct = bind.voxel_build(cfg)
    assert staging["rejected_bricks"] == 0 and direct["rejected_bricks"] > 0
    assert direct["active_bricks"] == staging["active_bricks"]
    assert direct["active_cells"] == staging["active_cells"]
    assert np.array_equal(_cells_through_pages(direct, 32), _cells_through_pages(staging, 32))


def _brick_distance_range(G, B, res, lo=-2.0, hi=2.0, radius=1.0):
    """Closed form of the sphere distance over each brick's box: min and max
    of |p| - r, (G^3,) arrays in linear brick order."""
    cell = (hi - lo) / res
    idx = np.arange(G)
    first = idx * B
    end = np.minimum(first + B, res)
    a0, a1 = lo + first * cell, lo + end * cell
    near = np.where(a1 < 0, a1, np.where(a0 > 0, a0, 0.0))          # per axis
    far = np.maximum(np.abs(a0), np.abs(a1))
    nz, ny, nx = np.meshgrid(near, near, near, indexing="ij")
    fz, fy, fx = np.meshgrid(far, far, far, indexing="ij")
    dmin = np.sqrt(nx ** 2 + ny ** 2 + nz ** 2) - radius
    dmax = np.sqrt(fx ** 2 + fy ** 2 + fz ** 2) - radius
    return dmin.reshape(-1), dmax.reshape(-1)


@pytest.mark.parametrize("storage,B", [(ffi.VOXEL_STORAGE_SPARSE_BRICKS, 4), (ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 8)])
@pytest.mark.parametrize("fill", [0, 1])
def test_rejection_drops_exactly_the_bricks_off_the_surface(storage, B, fill):
    """Legacy coverage with a wide band marks cells near, but not on, the
    surface.  Against the sphere's closed-form distance range over each brick
    box: a brick clearly outside (min > 1e-3), or clearly inside (max <
    -1e-3) with the interior not filled, keeps no page and is counted; a
    brick clearly straddling the surface keeps its cells; a brick inside
    with fill_interior keeps its filled cells."""
    res = 36
    cfg = _sphere_rejection_cfg(storage, ffi.VOXEL_COVERAGE_LEGACY, res=res, band=3.0, fill=fill)
    staging = bind.voxel_build(cfg)
    cfg.voxel_build_mode, cfg.voxel_brick_rejection = 1, 1
    direct = bind.voxel_build(cfg)
    G = -(-res // B)
    dmin, dmax = _brick_distance_range(G, B, res)
    outside, inside = dmin > 1e-3, dmax < -1e-3
    surface = (dmin < -1e-3) & (dmax > 1e-3)
    clear_reject = outside | (inside & (fill == 0))
    had = staging["page_table"] > 0
    assert np.all(direct["page_table"][clear_reject] == 0)
    assert np.all((direct["page_table"] > 0)[surface & had])
    if fill:
        assert np.all((direct["page_table"] > 0)[inside & had])
    ambiguous = ~(outside | inside | surface)
    assert clear_reject.sum() <= direct["rejected_bricks"] <= clear_reject.sum() + ambiguous.sum()
    # rejection dropped occupied cells here, unlike interval coverage
    assert (had & clear_reject).any() and direct["active_cells"] < staging["active_cells"]
    # kept bricks hold exactly the staging build's cells
    occ_s, occ_d = _cells_through_pages(staging, res), _cells_through_pages(direct, res)
    keep = np.zeros((G * B,) * 3, bool)
    keep.reshape(G, B, G, B, G, B)[...] = (direct["page_table"] > 0).reshape(G, 1, G, 1, G, 1)
    keep = keep[:res, :res, :res]
    assert np.array_equal(occ_d, occ_s & keep)


def test_direct_build_without_rejection_or_programs_equals_staging():
    cfg = _sphere_rejection_cfg(ffi.VOXEL_STORAGE_SPARSE_BRICKS, ffi.VOXEL_COVERAGE_LEGACY, band=3.0)
    staging = bind.voxel_build(cfg)
    cfg.voxel_build_mode = 1
    direct = bind.voxel_build(cfg)
    assert direct["rejected_bricks"] == 0 and np.array_equal(direct["page_table"], staging["page_table"])
    mb = _small(ffi.VOXEL_STORAGE_SPARSE_BRICKS)   # a built-in fractal: no certified bounds
    a = bind.voxel_build(mb)
    mb.voxel_build_mode, mb.voxel_brick_rejection = 1, 1
    b = bind.voxel_build(mb)
    assert b["rejected_bricks"] == 0 and np.array_equal(a["page_table"], b["page_table"])


@pytest.mark.parametrize("storage", [ffi.VOXEL_STORAGE_DENSE, ffi.VOXEL_STORAGE_SPARSE_BRICKS,
                                     ffi.VOXEL_STORAGE_SPARSE_BRICKS8])
def test_render_over_given_tables_equals_render_over_own_build(storage):
    """orc_voxel_render_tables (the CPU baseline over tables read back from
    the GPU) renders exactly what the oracle renders from its own build."""
    cfg = _small(storage, res=40)
    cfg.width, cfg.height, cfg.samples = 40, 24, 2
    own = bind.voxel_render(cfg)
    given = bind.voxel_render_tables(cfg, bind.voxel_build(cfg))
    assert np.array_equal(own, given)
