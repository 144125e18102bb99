"""GPU voxel backend against the FP32 CPU restatement.

Bit-exact (integer/byte work): page tables, brick pools (records, FP16
distances, u8 material ids), material palettes and DDA cell-visit sequences
for the same float32 rays.  Images: the camera rays use transcendentals
whose last ulp differs between the GPU and the host libm, so image parity is
a tolerance (RMSE <= 2e-2 at 2 spp, >= 90% of pixels within 1e-3).
"""
import numpy as np
import pytest

from fractalpathtracer_amd import api, ffi, scenes
from oracle import bind

pytestmark = pytest.mark.gpu


def test_reference_dda_contract_runs_on_gpu():
    api.test_voxel_dda()  # raises with the failure bitmask


def _cfg(storage, coverage, res):
    c = scenes.voxel_mandelbox_scene(96, 54, 2, resolution=res)
    c.voxel_storage = storage
    c.voxel_coverage_mode = coverage
    return c


LAYOUTS = [(ffi.VOXEL_STORAGE_DENSE, 0), (ffi.VOXEL_STORAGE_SPARSE_BRICKS, 0), (ffi.VOXEL_STORAGE_SPARSE_BRICKS, 3),
           (ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 0), (ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 3)]


@pytest.mark.parametrize("storage,coverage", LAYOUTS)
def test_tables_are_bit_exact(storage, coverage):
    cfg = _cfg(storage, coverage, 44)
    g = api.voxel_build(cfg)
    c = bind.voxel_build(cfg)
    for k in ("storage", "brick_size", "bricks_per_axis", "active_bricks", "active_cells"):
        assert g[k] == c[k], k
    assert np.array_equal(g["page_table"], c["page_table"])
    assert g["records"].tobytes() == c["records"].tobytes()
    assert np.array_equal(g["material"], c["material"])
    assert np.array_equal(g["distance"], c["distance"])
    if storage == ffi.VOXEL_STORAGE_SPARSE_BRICKS8:
        assert np.array_equal(g["palette"], c["palette"])


def _program_cfg(storage, coverage, res=40):
    """Box minus sphere through abs/scale/translate/sort (no transcendentals,
    so the distance and its certified bounds are IEEE-exact on both sides)."""
    c = scenes.voxel_mandelbox_scene(64, 48, 1, resolution=res)
    scenes.set_program(c, [(ffi.OP_ABS, 0, []), (ffi.OP_SCALE, 0, [1.25]), (ffi.OP_TRANSLATE, 0, [0.2, -0.1, 0.3]),
                           (ffi.OP_SORT_DESC, 0, []), (ffi.OP_BOX, 0, [1.0, 0.8, 0.6]), (ffi.OP_SPHERE, 2, [0.9])])
    for i, v in enumerate([-2.0, -2.0, -2.0]):
        c.voxel_bounds_min[i] = v
    for i, v in enumerate([2.0, 2.0, 2.0]):
        c.voxel_bounds_max[i] = v
    c.voxel_storage = storage
    c.voxel_coverage_mode = coverage
    return c


@pytest.mark.parametrize("storage", [ffi.VOXEL_STORAGE_DENSE, ffi.VOXEL_STORAGE_SPARSE_BRICKS,
                                     ffi.VOXEL_STORAGE_SPARSE_BRICKS8])
@pytest.mark.parametrize("coverage", [ffi.VOXEL_COVERAGE_LEGACY, ffi.VOXEL_COVERAGE_LIPSCHITZ,
                                      ffi.VOXEL_COVERAGE_INTERVAL])
def test_program_coverage_tables_are_bit_exact(storage, coverage):
    cfg = _program_cfg(storage, coverage)
    g = api.voxel_build(cfg)
    c = bind.voxel_build(cfg)
    print("coverage", coverage, "cells", g["active_cells"], "bricks", g["active_bricks"])
    for k in ("active_bricks", "active_cells"):
        assert g[k] == c[k], k
    assert np.array_equal(g["page_table"], c["page_table"])
    assert g["records"].tobytes() == c["records"].tobytes()
    assert np.array_equal(g["material"], c["material"])
    assert np.array_equal(g["distance"], c["distance"])


def test_tables_bit_exact_at_256_cubed():
    cfg = _cfg(ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 3, 256)
    g = api.voxel_build(cfg)
    c = bind.voxel_build(cfg)
    print("256^3 shell: bricks %d/%d cells %d, GPU build %.1f ms" % (
        g["active_bricks"], g["bricks_per_axis"] ** 3, g["active_cells"], g["build_ms"]))
    assert np.array_equal(g["page_table"], c["page_table"])
    assert np.array_equal(g["material"], c["material"])
    assert np.array_equal(g["distance"], c["distance"])


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.normal(0, 1, (n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * rng.uniform(6, 14, (n, 1))
    d = rng.uniform(-5, 5, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # include axis-aligned and zero-component directions (ties, slabs)
    d[:40] = np.round(d[:40])
    d[:40] /= np.maximum(np.linalg.norm(d[:40], axis=1, keepdims=True), 1e-9)
    return np.concatenate([o, d], axis=1).astype(np.float32)


@pytest.mark.parametrize("storage,coverage", LAYOUTS)
def test_cell_visit_sequences_are_bit_exact(storage, coverage):
    cfg = _cfg(storage, coverage, 64)
    rays = _rays(1500, storage * 7 + coverage)
    gc, gn, gh = api.voxel_trace_cells(cfg, rays, cap=400)
    cc, cn, ch = bind.voxel_trace_cells(cfg, rays, cap=400)
    assert np.array_equal(gn, cn)
    for i in range(len(rays)):
        k = min(int(cn[i]), 400)
        assert np.array_equal(gc[i, :k], cc[i, :k]), i
    assert gh.tobytes() == ch.tobytes()
    assert gh[:, 0].sum() > 100


@pytest.mark.parametrize("storage", [ffi.VOXEL_STORAGE_SPARSE_BRICKS, ffi.VOXEL_STORAGE_SPARSE_BRICKS8])
def test_voxel_render_matches_fp32_restatement(storage):
    cfg = _cfg(storage, 3, 64)
    cfg.renderer_backend = ffi.RENDERER_VOXEL
    res = api.render_full(cfg)
    ref = bind.voxel_render(cfg)
    d = np.abs(res.linear[..., :3] - ref[..., :3])
    rmse = float(np.sqrt(np.mean(d ** 2)))
    close = float(np.mean(d.max(axis=-1) <= 1e-3))
    print("voxel image rmse %.3e, pixels within 1e-3: %.4f, active cells %d" % (rmse, close, res.voxel["active_cells"]))
    assert rmse <= 2e-2 and close >= 0.9
    assert res.voxel["active_cells"] > 0 and res.build_ms > 0


def _refined(mode):
    """Secant/bisection and restricted trace on the typed program (box minus
    sphere), fixed DE on the built-in Menger sponge (initial cut scale 1,
    divisor 3: its distance is tight enough for the method's acceptance
    bound; the Mandelbox's conservative estimate rarely is)."""
    if mode == ffi.VOXEL_LEAF_REFINEMENT_FIXED_DE:
        c = _cfg(ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 0, 128)
        c.sdf_id = ffi.SDF_MENGER_SPONGE
        c.set_values[0], c.set_values[1], c.set_values[2], c.set_values[3] = 1.0, 3.0, 0.0, 0.0
        for i in range(3):
            c.voxel_bounds_min[i], c.voxel_bounds_max[i] = -1.2, 1.2
        c.camera_position[0], c.camera_position[1], c.camera_position[2] = 0.0, 0.0, -3.5
        c.camera_yaw_pitch[0], c.camera_yaw_pitch[1] = 0.5, 0.4
    else:
        c = _program_cfg(ffi.VOXEL_STORAGE_SPARSE_BRICKS8, 0, 48)
        c.camera_position[2] = -3.5
        c.camera_fov = 50.0
    c.renderer_backend = ffi.RENDERER_VOXEL
    c.voxel_leaf_refinement = mode
    return c


@pytest.mark.parametrize("mode", [ffi.VOXEL_LEAF_REFINEMENT_SECANT_BISECTION,
                                  ffi.VOXEL_LEAF_REFINEMENT_RESTRICTED_TRACE,
                                  ffi.VOXEL_LEAF_REFINEMENT_FIXED_DE])
def test_leaf_refinement_render_and_depth_match_oracle(mode):
    cfg = _refined(mode)
    res = api.render_full(cfg)
    ref = bind.voxel_render(cfg)
    d = np.abs(res.linear[..., :3] - ref[..., :3])
    rmse = float(np.sqrt(np.mean(d ** 2)))
    close = float(np.mean(d.max(axis=-1) <= 1e-3))
    plain = scenes.clone(cfg)
    plain.voxel_leaf_refinement = ffi.VOXEL_LEAF_REFINEMENT_NONE
    changed = float(np.mean(np.any(api.render_full(plain).linear != res.linear, axis=-1)))
    img, _ = api.diagnostic(cfg, "depth", max_distance=30.0)
    dref = bind.diagnostic(cfg, ffi.DIAGNOSTIC_MODES["depth"], max_distance=30.0)
    dclose = float(np.mean(np.abs(img.astype(int) - dref.astype(int)) <= 1))
    print(mode, "rmse %.3e within 1e-3 %.4f; pixels changed by refinement %.3f; depth within 1: %.4f" % (
        rmse, close, changed, dclose))
    assert rmse <= 2e-2 and close >= 0.9
    assert changed > 0.01
    assert dclose >= 0.99


@pytest.mark.parametrize("storage", [ffi.VOXEL_STORAGE_SPARSE_BRICKS, ffi.VOXEL_STORAGE_SPARSE_BRICKS8])
def test_trace_memory_counts_match_oracle(storage):
    """fpt_hip_plan_voxel_stats against the FP32 restatement's counts of the
    same frame: within 1% (the camera rays' last-ulp differences move a few
    traversals, as in the image tolerance above)."""
    cfg = _cfg(storage, 3, 64)
    cfg.renderer_backend = ffi.RENDERER_VOXEL
    with api.Plan(cfg) as plan:
        buf = api.DeviceBuffer(cfg.width * cfg.height * 4)
        st = plan.voxel_stats(buf)
        img = buf.numpy().reshape(cfg.height, cfg.width, 4)
    ref = bind.voxel_render(cfg)
    k = bind.voxel_counts()
    print(storage, st, k)
    assert st["pixels"] == cfg.width * cfg.height
    assert float(np.sqrt(np.mean((img[..., :3] - ref[..., :3]) ** 2))) <= 2e-2
    for key in ("visited_cells", "page_reads", "cell_reads"):
        assert k[key] > 0 and abs(st[key] - k[key]) <= 0.01 * k[key], (key, st[key], k[key])


@pytest.mark.parametrize("storage", [ffi.VOXEL_STORAGE_SPARSE_BRICKS, ffi.VOXEL_STORAGE_SPARSE_BRICKS8])
@pytest.mark.parametrize("coverage,fill", [(ffi.VOXEL_COVERAGE_LEGACY, 0), (ffi.VOXEL_COVERAGE_LEGACY, 1),
                                           (ffi.VOXEL_COVERAGE_INTERVAL, 0), (ffi.VOXEL_COVERAGE_LIPSCHITZ, 0)])
def test_brick_rejection_tables_are_bit_exact(storage, coverage, fill):
    """The direct build with certified whole-brick rejection
    (Shaders.metal:5704-5726): tables and rejected-brick counts equal the FP32
    oracle's, on the box-minus-sphere program with a wide legacy band (where
    rejection drops occupied cells)."""
    cfg = _program_cfg(storage, coverage, res=44)
    cfg.voxel_surface_band = 3.0
    cfg.voxel_fill_interior = fill
    cfg.voxel_build_mode, cfg.voxel_brick_rejection = 1, 1
    g = api.voxel_build(cfg)
    c = bind.voxel_build(cfg)
    print("coverage", coverage, "fill", fill, "bricks", g["active_bricks"], "rejected", g["rejected_bricks"])
    assert g["rejected_bricks"] == c["rejected_bricks"] > 0
    for k in ("active_bricks", "active_cells"):
        assert g[k] == c[k], k
    assert np.array_equal(g["page_table"], c["page_table"])
    assert g["records"].tobytes() == c["records"].tobytes()
    assert np.array_equal(g["material"], c["material"])
    assert np.array_equal(g["distance"], c["distance"])


def test_direct_build_refuses_dense_storage():
    cfg = _program_cfg(ffi.VOXEL_STORAGE_DENSE, ffi.VOXEL_COVERAGE_LEGACY)
    cfg.voxel_build_mode = 1
    with pytest.raises(ffi.FptError, match="direct voxel build requires a sparse voxel storage mode"):
        api.voxel_build(cfg)


def test_metal_render_reports_rejected_bricks():
    cfg = _program_cfg(ffi.VOXEL_STORAGE_SPARSE_BRICKS, ffi.VOXEL_COVERAGE_LEGACY, res=44)
    cfg.renderer_backend = ffi.RENDERER_VOXEL
    cfg.voxel_surface_band = 3.0
    cfg.width, cfg.height, cfg.samples = 32, 24, 1
    cfg.voxel_build_mode, cfg.voxel_brick_rejection = 1, 1
    res = api.render_full(cfg)
    assert res.voxel["rejected_bricks"] == bind.voxel_build(cfg)["rejected_bricks"] > 0


@pytest.mark.parametrize("storage", [ffi.VOXEL_STORAGE_SPARSE_BRICKS, ffi.VOXEL_STORAGE_SPARSE_BRICKS8])
@pytest.mark.parametrize("fill", [0, 1])
def test_brick_rejection_on_the_sphere_is_bit_exact(storage, fill):
    """The sphere scene has bricks certified inside, so the fill_interior
    rule decides their fate (rejected when the interior is not filled)."""
    from tests.test_voxel_oracle import _sphere_rejection_cfg
    cfg = _sphere_rejection_cfg(storage, ffi.VOXEL_COVERAGE_LEGACY, res=36, band=3.0, fill=fill)
    cfg.voxel_build_mode, cfg.voxel_brick_rejection = 1, 1
    g = api.voxel_build(cfg)
    c = bind.voxel_build(cfg)
    print("fill", fill, "bricks", g["active_bricks"], "rejected", g["rejected_bricks"])
    assert g["rejected_bricks"] == c["rejected_bricks"] > 0
    assert (g["active_bricks"], g["active_cells"]) == (c["active_bricks"], c["active_cells"])
    assert np.array_equal(g["page_table"], c["page_table"])
    assert g["records"].tobytes() == c["records"].tobytes()
    assert np.array_equal(g["material"], c["material"])


def _voxel_frame(cfg, kernel, **tune):
    with api.Plan(cfg) as plan:
        plan.tune(kernel=kernel, **tune)
        buf = api.DeviceBuffer(cfg.width * cfg.height * 4)
        plan.render_into(buf.address)
        img = buf.numpy().reshape(cfg.height, cfg.width, 4).copy()
        stats = plan.voxel_stats(buf)
    return img, stats


@pytest.mark.parametrize("storage,coverage", LAYOUTS)
@pytest.mark.parametrize("scene,spp,accumulation", [("mandelbox", 3, ffi.SDF_ACCUMULATION_BATCH),
                                                    ("program", 5, ffi.SDF_ACCUMULATION_CHUNKED)])
def test_persistent_voxel_kernel_matches_the_one_path_kernel(storage, coverage, scene, spp, accumulation):
    """The path-regenerating voxel kernel (lane state machine, batched
    shading, pixels from a global counter) renders exactly the image of the
    one-pixel-per-lane kernel, with the same traversal counters, for every
    storage layout; with chunked accumulation every launch resumes the
    running mean from the buffer.  The program scene's material is
    translucent and emissive, so the refraction and emission branches run
    beside the sun rays."""
    if scene == "mandelbox":
        cfg = _cfg(storage, coverage, 48)
    else:
        cfg = _program_cfg(storage, coverage)
        cfg.program_material[5] = 0.3   # translucency
        cfg.program_material[7] = 0.4   # emission
    cfg.width, cfg.height, cfg.samples = 77, 45, spp
    cfg.sdf_accumulation_mode = accumulation
    cfg.sdf_chunk_samples = 2
    a, sa = _voxel_frame(cfg, 0)
    b, sb = _voxel_frame(cfg, 1)
    c, sc = _voxel_frame(cfg, 1, workgroups=7, event_batch=1)   # few lanes: many slots per lane
    assert np.isfinite(a).all() and a[..., :3].max() > 0
    assert np.array_equal(a, b) and np.array_equal(a, c)
    assert sa == sb == sc
