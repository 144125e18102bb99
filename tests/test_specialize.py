"""Topology specialisation (runtime-compiled straight-line programs,
metal_bridge.mm:592-1300, 4469-4500; validation Shaders.metal:5055-5103).

CPU: the generated source has one straight-line step per instruction and
rejects what the interpreter cannot run.  GPU: the specialised kernel is
validated on 2^20 points, renders bit-identically to the interpreter (the
same per-instruction steps in the same order), is served from the memory
and disk caches, and reports through fpt_metal_render's stitch outputs."""
import os
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from fractalpathtracer_amd import api, ffi, scenes

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"


def _csg():
    return scenes.apply_scene_json(scenes.default_config(), GOLDEN / "scenes" / "Exact_CSG.json")


def test_generated_source_is_straight_line():
    c = scenes.typed_program_scene(32, 32, 1)
    src = api.topology_source(c)
    n = c.sdf_program_count
    steps = re.findall(r"eval_op<(\d+)u, false>\(ops\[(\d+)\]", src)
    assert [int(i) for _, i in steps] == list(range(n))
    assert [int(o) for o, _ in steps] == [c.sdf_program[i].opcode for i in range(n)]
    surf = re.findall(r"surf_op<(\d+)u>\(ops\[(\d+)\]", src)
    assert len(surf) == n and "#define FPT_TOPOLOGY_SPECIALIZED 1" in src
    assert "fpt_topology_validate" in src and "for (" not in src.split("topology_distance")[1].split("}")[0]


def test_generated_source_errors():
    c = scenes.mandelbulb_scene(8, 8, 1)
    c.sdf_id = ffi.SDF_MANDELBOX_FRACTAL
    with pytest.raises(ffi.FptError, match="typed geometry program"):
        api.topology_source(c)
    c = _csg()
    c.sdf_program[0].opcode = 77
    with pytest.raises(ffi.FptError, match="unsupported opcode 77"):
        api.topology_source(c)


# --------------------------------------------------------------------------
# GPU.
# --------------------------------------------------------------------------

def _render(cfg, specialize):
    with api.Plan(cfg) as plan:
        stats = plan.specialize() if specialize else None
        buf = api.DeviceBuffer(cfg.width * cfg.height * 4)
        plan.render_into(buf.address)
        return buf.numpy().reshape(cfg.height, cfg.width, 4), stats


@pytest.mark.gpu
@pytest.mark.parametrize("scene", ["typed", "csg", "fold", "transforms", "union"])
def test_specialized_render_is_bit_identical(scene):
    """"transforms" is the fixture of the reference's topology validator
    (main.rs:4338-4385, Exact_Transforms.json)."""
    fixture = {"fold": "Typed_Program_Fold", "transforms": "Exact_Transforms", "union": "Exact_Translation_Union"}
    if scene in fixture:
        cfg = scenes.apply_scene_json(scenes.default_config(), GOLDEN / "scenes" / f"{fixture[scene]}.json")
    else:
        cfg = {"typed": lambda: scenes.typed_program_scene(96, 64, 3), "csg": lambda: _csg()}[scene]()
    cfg.width, cfg.height, cfg.samples = 96, 64, 3
    ref, _ = _render(cfg, False)
    img, st = _render(cfg, True)
    print(scene, st)
    assert st["sample_count"] == 1 << 20 and st["distance_failures"] == 0 and st["gradient_failures"] == 0
    assert st["cache_status"] in (1, 2) and st["code_bytes"] > 0
    assert np.array_equal(img, ref)


@pytest.mark.gpu
def test_module_cache_serves_same_topology_with_new_parameters():
    a = scenes.typed_program_scene(32, 32, 1)
    b = scenes.typed_program_scene(32, 32, 1)
    b.sdf_program[1].data[0] = -1.7       # new numbers, same opcodes
    with api.Plan(a) as pa:
        pa.specialize()
    with api.Plan(b) as pb:
        st = pb.specialize()
    assert st["cache_status"] == 2 and st["compile_ms"] < 1000 and st["distance_failures"] == 0


@pytest.mark.gpu
def test_disk_cache_in_a_fresh_process(tmp_path):
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from fractalpathtracer_amd import api, scenes\n"
            "c = scenes.typed_program_scene(16, 16, 1)\n"
            "with api.Plan(c) as p:\n"
            "    print(p.specialize()['cache_status'])\n" % str(ROOT))
    env = dict(os.environ, FPT_HIPRTC_CACHE=str(tmp_path))
    first = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    second = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert first.returncode == 0 and second.returncode == 0, first.stderr[-1500:] + second.stderr[-1500:]
    assert first.stdout.split()[-1] == "1" and second.stdout.split()[-1] == "2"
    assert len(list(tmp_path.glob("topology-*.co"))) == 1


@pytest.mark.gpu
def test_metal_render_honours_topology_request_and_reports_stitch_stats():
    cfg = scenes.typed_program_scene(64, 48, 2)
    plain = api.render_full(cfg)
    cfg.sdf_topology_specialization = 1
    res = api.render_full(cfg)
    st = res.stitch
    assert st["cache_status"] in (1, 2)
    assert st["validation"]["sample_count"] == 1 << 20
    assert st["validation"]["distance_failures"] == 0 and st["validation"]["gradient_failures"] == 0
    p = st["pipeline"]
    assert p["instruction_count"] == cfg.sdf_program_count and p["primitive_instruction_count"] == 2
    assert p["runtime_source_bytes"] > 0 and p["thread_execution_width"] == 64
    assert np.array_equal(res.linear, plain.linear)


@pytest.mark.gpu
def test_specialization_errors():
    vox = scenes.voxel_mandelbox_scene(16, 16, 1, resolution=32)
    scenes.set_program(vox, [(ffi.OP_SPHERE, 0, [1.0])])
    with api.Plan(vox) as p:
        with pytest.raises(ffi.FptError, match="direct SDF renderer"):
            p.specialize()
    with api.Plan(scenes.voxel_mandelbox_scene(16, 16, 1, resolution=32)) as p:
        with pytest.raises(ffi.FptError):
            p.specialize()


def _topology_cfg(scene, w=96, h=64, spp=3):
    fixture = {"fold": "Typed_Program_Fold", "transforms": "Exact_Transforms", "union": "Exact_Translation_Union"}
    if scene in fixture:
        cfg = scenes.apply_scene_json(scenes.default_config(), GOLDEN / "scenes" / f"{fixture[scene]}.json")
    else:
        cfg = {"typed": lambda: scenes.typed_program_scene(96, 64, 3), "csg": lambda: _csg()}[scene]()
    cfg.width, cfg.height, cfg.samples = w, h, spp
    return cfg


@pytest.mark.gpu
@pytest.mark.parametrize("scene", ["typed", "csg", "fold", "transforms", "union"])
def test_specialized_code_against_the_fp64_oracle(scene):
    """The specialised distance and gradient themselves, at 4096 points of
    the scene's bounds (half of them pulled close to the surface), against the
    FP64 oracle's program rather than the interpreter:
    |d - d64| <= 5e-5 (1 + |d64|) everywhere (main.rs:3297-3361's
    tolerance), and the normalised gradient within 1e-4 of the FP64 one on
    at least as many points as the oracle's own FP32 twin manages (the
    misses are creases of folds, CSG and fractal leaves, where the float32
    rounding of the point picks another branch).  Then the specialised
    render against the FP64 image (the image bar)."""
    from oracle import bind
    cfg = _topology_cfg(scene)
    rng = np.random.default_rng(11)
    lo, hi = np.array(cfg.voxel_bounds_min), np.array(cfg.voxel_bounds_max)
    pts = rng.uniform(lo, hi, (4096, 3)).astype(np.float32)
    s64 = bind.program_surface(cfg, pts.astype(np.float64), 64)
    g = s64[:, 1:4] / np.maximum(np.linalg.norm(s64[:, 1:4], axis=1, keepdims=True), 1e-30)
    # Half of the points moved to 3% of their distance from the surface: an
    # exact projection would land on the creases of boxes (edges, corners),
    # where the gradient has no direction to compare.
    pts[2048:] = (pts.astype(np.float64) - g * s64[:, :1] * 0.97).astype(np.float32)[2048:]
    ref = bind.program_surface(cfg, pts.astype(np.float64), 64)
    twin = bind.program_surface(cfg, pts.astype(np.float64), 32)
    with api.Plan(cfg) as plan:
        plan.specialize()
        got = plan.specialized_surface(pts)
    d_err = np.abs(got[:, 0] - ref[:, 0])
    assert np.all(np.isfinite(got)) and np.all(d_err <= 5e-5 * (1 + np.abs(ref[:, 0]))), d_err.max()

    def unit(v):
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    rn = unit(ref[:, 1:4])
    frac = float(np.mean(np.abs(unit(got[:, 1:4]) - rn).max(axis=1) <= 1e-4))
    frac_twin = float(np.mean(np.abs(unit(twin[:, 1:4]) - rn).max(axis=1) <= 1e-4))
    print(scene, f"distance max err {d_err.max():.2e}; gradient within 1e-4 of FP64 on {frac:.4f} "
                 f"(FP32 twin {frac_twin:.4f})")
    assert frac >= frac_twin - 0.005 and frac >= 0.95
    img, _ = _render(cfg, True)
    img64 = bind.render(cfg, 64)
    img32 = bind.render(cfg, 32)
    rmse = float(np.sqrt(np.mean((img[..., :3] - img64[..., :3]) ** 2)))
    rmse_twin = float(np.sqrt(np.mean((img32[..., :3] - img64[..., :3]) ** 2)))
    print(scene, f"specialised image rmse {rmse:.3e} twin {rmse_twin:.3e}")
    assert rmse <= 1.25 * rmse_twin + 2e-3
