"""Direct evaluators of lowered programs on the CPU: the flat-union,
canonical affine and typed structure-of-arrays tables that the scene front
end writes (scene.rs:1345-1530) and that the renderer evaluates in place of
the instruction interpreter (Shaders.metal:2713-2970, dispatch 3576-3584 for
distances and 3403-3408 for analytic surface gradients).

Pins: the oracle's table evaluators against its own interpreter of the same
program (program_eval / program_surface, pinned elsewhere by closed forms),
on every bundled scene that each lowering accepts.  The lowerings are exact
for these programs (translations, unit-scale affines, hard unions and the
canonical CSG list), so in FP64 distances agree bit for bit, primary hits
agree and their positions and surface normals agree within 1e-12 (the
interpreter applies +1.2 then -1.2 to a point in double where the table
holds their exact float sum, 0, and the canonical gradient is a product
with the affine rows); in the FP32 twin an affine row applied to a point
rounds differently from a subtracted translation, so distances agree within
4e-7 (1 + |d|) (about 2 ulp) and normals within 1e-6.  A tie between sources with
different gradients pins the typed table's rule that the earlier program
instruction wins (Shaders.metal:2898-2970).
"""
import json

import numpy as np
import pytest

from fractalpathtracer_amd import ffi, scenes
from oracle import bind
from test_scene_front import GOLDEN, _scene_tree, _write_preset

LOWERED = [
    ("Exact_Box", "flat_union"), ("Exact_Box", "canonical_ir"), ("Exact_Box", "typed_soa"),
    ("Exact_CSG", "canonical_ir"),
    ("Exact_Plane", "flat_union"), ("Exact_Plane", "canonical_ir"), ("Exact_Plane", "typed_soa"),
    ("Exact_Sphere", "flat_union"), ("Exact_Sphere", "canonical_ir"), ("Exact_Sphere", "typed_soa"),
    ("Exact_Tie", "canonical_ir"),
    ("Exact_Translation_Union", "flat_union"), ("Exact_Translation_Union", "canonical_ir"),
    ("Exact_Translation_Union", "typed_soa"),
]


def _counts(c):
    t = c.sdf_typed_soa
    return c.sdf_flat_union_count, c.sdf_canonical_count, t.sphere_count + t.box_count + t.plane_count


def interpreted(c):
    """The same config with every lowered table dropped: the interpreter runs."""
    d = scenes.clone(c)
    d.sdf_flat_union_count = d.sdf_canonical_count = 0
    d.sdf_typed_soa.sphere_count = d.sdf_typed_soa.box_count = d.sdf_typed_soa.plane_count = 0
    return d


def lowered_scene(tmp_path, name, option):
    path, _ = _scene_tree(tmp_path, name)
    cfg, _ = scenes.load_scene(path, fpt_root=tmp_path, **{option: 1})
    return cfg


@pytest.mark.parametrize("name,option", LOWERED)
def test_lowered_tables_match_the_interpreter(tmp_path, name, option):
    c = lowered_scene(tmp_path, name, option)
    fu, can, soa = _counts(c)
    assert {"flat_union": fu, "canonical_ir": can, "typed_soa": soa}[option] > 0
    assert sum(x > 0 for x in (fu, can, soa)) == 1
    i = interpreted(c)
    c.width, c.height, i.width, i.height = 40, 24, 40, 24
    pts = np.random.default_rng(7).uniform(-4, 4, (3000, 3))
    assert np.array_equal(bind.map_distance(c, pts, 64), bind.map_distance(i, pts, 64))
    d32_l, d32_i = bind.map_distance(c, pts, 32), bind.map_distance(i, pts, 32)
    assert np.all(np.abs(d32_l - d32_i) <= 4e-7 * (1 + np.abs(d32_i)))
    tr_l, tr_i = bind.trace_primary(c, 64), bind.trace_primary(i, 64)
    assert np.array_equal(tr_l[..., 7], tr_i[..., 7])
    assert np.all(np.abs(tr_l - tr_i) <= 1e-12 * (1 + np.abs(tr_i)))
    hit = tr_i[..., 7] == 1
    assert hit.any()
    surf = tr_i[..., 1:4][hit]
    assert np.abs(bind.normal(c, surf, 64) - bind.normal(i, surf, 64)).max() <= 1e-12
    assert np.abs(bind.normal(c, surf, 32) - bind.normal(i, surf, 32)).max() <= 1e-6


def test_typed_table_tie_goes_to_the_earlier_instruction(tmp_path):
    """Plane z = 0 (instruction 0) and a unit sphere (instruction 1) are both
    at distance 4 from (3, 0, 4); the table visits spheres before planes, so
    only the source-order rule makes the plane's gradient win, as the
    interpreter's first-wins union does."""
    src = json.loads((GOLDEN / "scenes" / "Exact_Sphere.json").read_text())
    src["sdf_program"]["operations"] = [{"op": "plane", "data": [0.0, 0.0, 1.0, 0.0]},
                                        {"op": "sphere", "radius": 1.0}]
    dst = tmp_path / "benchmarks" / "tie.json"
    dst.parent.mkdir(parents=True)
    dst.write_text(json.dumps(src))
    _write_preset((dst.parent / src["preset"]).resolve(), sets=[0.25, -1.5, 3.0])
    c, _ = scenes.load_scene(dst, fpt_root=tmp_path, typed_soa=1)
    t = c.sdf_typed_soa
    assert (t.sphere_count, t.plane_count, t.sphere_source[0], t.plane_source[0]) == (1, 1, 1, 0)
    p = np.array([[3.0, 0.0, 4.0]])
    c.sdf_normal_mode = 0
    for prec in (64, 32):
        assert bind.map_distance(c, p, prec)[0] == 4.0
        assert np.array_equal(bind.normal(c, p, prec)[0], [0.0, 0.0, 1.0])
        assert np.array_equal(bind.normal(interpreted(c), p, prec)[0], [0.0, 0.0, 1.0])
    # with the plane listed as the later instruction the sphere wins instead
    t.sphere_source[0], t.plane_source[0] = 0, 1
    assert np.allclose(bind.normal(c, p, 64)[0], [0.6, 0.0, 0.8], atol=1e-12)


def test_canonical_csg_list_evaluates_intersection_and_subtraction():
    """A canonical list [sphere r=1, box half 0.5 intersect, sphere r=0.25
    subtract] equals max(max(s1, box), -s2) at random points: the three
    combine codes of the canonical list (Shaders.metal:2751-2770)."""
    c = scenes.default_config()
    scenes.set_program(c, [(ffi.OP_SPHERE, 0, (1.0,))])
    ops = [(ffi.OP_SPHERE, (1.0, 0, 0, 0), 0), (ffi.OP_BOX, (0.5, 0.5, 0.5, 0), 1),
           (ffi.OP_SPHERE, (0.25, 0, 0, 0), 2)]
    for k, (op, data, combine) in enumerate(ops):
        inst = c.sdf_canonical_primitives[k]
        inst.opcode, inst._pad0, inst.distance_scale = op, combine, 1.0
        for j, v in enumerate(data):
            inst.data[j] = v
        for j, v in enumerate((1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0)):
            inst.transform[j] = v
    c.sdf_canonical_count = 3
    p = np.random.default_rng(3).uniform(-1.5, 1.5, (2000, 3))
    s1 = np.linalg.norm(p, axis=1) - 1.0
    q = np.abs(p) - 0.5
    box = np.minimum(q.max(axis=1), 0) + np.linalg.norm(np.maximum(q, 0), axis=1)
    s2 = np.linalg.norm(p, axis=1) - 0.25
    want = np.maximum(np.maximum(s1, box), -s2)
    assert np.allclose(bind.map_distance(c, p, 64), want, rtol=0, atol=1e-6)


def test_typed_soa_validator_refuses_configs_without_a_typed_table(tmp_path):
    """fpt_test_typed_soa's input checks (metal_bridge.mm:4122-4134) run
    before any device work."""
    from fractalpathtracer_amd import api
    c = lowered_scene(tmp_path, "Exact_Translation_Union", "canonical_ir")
    with pytest.raises(RuntimeError, match="requires a lowered typed program"):
        api.test_typed_soa(c)
    m = scenes.mandelbulb_scene(8, 8)
    with pytest.raises(RuntimeError, match="requires a lowered typed program"):
        api.test_typed_soa(m)


def test_primitive_union_workload_lowers_to_a_28_entry_table():
    """The lowered-program workload: the library's pass turns the 55-
    instruction program into 15 sphere, 12 box and 1 plane entries, whose
    distances match the interpreter's (the table holds float sums of the
    translations, the interpreter applies them one by one)."""
    c = scenes.primitive_union_scene(64, 36, 1)
    t = c.sdf_typed_soa
    assert (c.sdf_program_count, t.sphere_count, t.box_count, t.plane_count) == (55, 15, 12, 1)
    assert _counts(scenes.primitive_union_scene(64, 36, 1, lowered=False)) == (0, 0, 0)
    i = interpreted(c)
    pts = np.random.default_rng(2).uniform(-5, 6, (4000, 3))
    d_l, d_i = bind.map_distance(c, pts, 64), bind.map_distance(i, pts, 64)
    assert np.all(np.abs(d_l - d_i) <= 1e-6 * (1 + np.abs(d_i)))
    tr = bind.trace_primary(c, 64)
    assert np.array_equal(tr[..., 7], bind.trace_primary(i, 64)[..., 7]) and tr[..., 7].all()
