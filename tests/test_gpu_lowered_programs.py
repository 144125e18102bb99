"""Lowered programs on the GPU: scenes whose program the front end lowered
to a flat-union, canonical affine or typed structure-of-arrays table render
through the table evaluators (Shaders.metal:2713-2970, dispatch 3576-3584
and 3403-3408), checked against the oracle's evaluators of the same tables.

Bars, as for the other SDF scenes:
  field          distance within 5e-5 (1 + |d|) of FP64 at every point;
  primary rays   hit masks agree on >= 99.9% of pixels, t within 1e-4
                 relative on >= 99.5% of common hits (p99 <= 1e-5), normals
                 within 1e-4 on >= 98.5%;
  images         RMSE to the FP64 image at most 1.25x the FP32 twin's +
                 2e-3;
  kernels        the persistent and one-path-per-lane kernels render
                 bit-identical frames; a batch renders each scene as it
                 renders alone; the learned march with a proposer below the
                 switch distance renders the exact image;
  validator      fpt_test_typed_soa's failure count matches the oracle's at
                 the same hashed points within 0.2%.
  voxels         tables built from lowered scenes are bit-exact against the
                 FP32 oracle.
"""
import numpy as np
import pytest

from fractalpathtracer_amd import api, scenes
from oracle import bind
from test_lowered_programs import interpreted, lowered_scene

pytestmark = pytest.mark.gpu

CASES = [("Exact_Translation_Union", "flat_union"), ("Exact_Translation_Union", "canonical_ir"),
         ("Exact_Translation_Union", "typed_soa"), ("Exact_CSG", "canonical_ir"),
         ("Exact_Tie", "canonical_ir"), ("Exact_Box", "typed_soa"), ("Exact_Plane", "typed_soa")]


@pytest.mark.parametrize("name,option", CASES)
def test_field_and_primary_rays_match_oracle(tmp_path, name, option):
    c = lowered_scene(tmp_path, name, option)
    c.width, c.height = 96, 54
    pts = np.random.default_rng(5).uniform(-4, 4, (8192, 3))
    got = api.sample_field(c, pts)[:, 0]
    want = bind.map_distance(c, pts.astype(np.float32).astype(np.float64), 64)
    assert np.all(np.abs(got - want) <= 5e-5 * (1 + np.abs(want)))
    # the same config without its table runs the interpreter on the GPU
    interp = api.sample_field(interpreted(c), pts)[:, 0]
    assert np.all(np.abs(got - interp) <= 4e-7 * (1 + np.abs(interp)))

    h = api.trace_primary(c)
    r = bind.trace_primary(c, 64)
    hit_g, hit_r = h["hit"] == 1, r[..., 7] == 1
    assert np.mean(hit_g == hit_r) >= 0.999
    both = hit_g & hit_r
    assert both.sum() > 20
    t_err = np.abs(h["t"][both] - r[..., 0][both]) / r[..., 0][both]
    print(name, option, "t frac<=1e-4 %.5f p99 %.2e" % (np.mean(t_err <= 1e-4), np.percentile(t_err, 99)))
    assert np.mean(t_err <= 1e-4) >= 0.995 and np.percentile(t_err, 99) <= 1e-5
    n64 = bind.normal(c, h["position"][both].astype(np.float64))
    n_err = np.linalg.norm(h["normal"][both] - n64, axis=-1)
    assert np.mean(n_err <= 1e-4) >= 0.985


@pytest.mark.parametrize("name,option", [("Exact_Translation_Union", "canonical_ir"),
                                         ("Exact_Translation_Union", "typed_soa"), ("Exact_CSG", "canonical_ir")])
def test_image_parity_and_kernel_agreement(tmp_path, name, option):
    c = lowered_scene(tmp_path, name, option)
    c.width, c.height, c.samples = 64, 36, 2
    img = api.render(c)
    ref = bind.render(c, 64)
    twin = bind.render(c, 32)
    rmse = float(np.sqrt(np.mean((img[..., :3] - ref[..., :3]) ** 2)))
    rmse_twin = float(np.sqrt(np.mean((twin[..., :3] - ref[..., :3]) ** 2)))
    print(name, option, "rmse gpu %.3e twin %.3e" % (rmse, rmse_twin))
    assert np.all(np.isfinite(img)) and img[..., :3].mean() > 0.01
    assert rmse <= 1.25 * rmse_twin + 2e-3
    frames = []
    for kernel in (1, 0):
        with api.Plan(c) as plan:
            plan.tune(kernel=kernel)
            buf = api.DeviceBuffer(c.width * c.height * 4)
            plan.render_into(buf.address)
            frames.append(buf.numpy())
    assert np.array_equal(frames[0], frames[1])


def test_batch_with_lowered_scenes_equals_single_renders(tmp_path):
    """A batch that holds lowered scenes runs the batched kernel's instance
    with the table evaluators; every scene, lowered or not, renders exactly
    as it does alone."""
    cfgs = [lowered_scene(tmp_path, "Exact_Translation_Union", "canonical_ir"),
            lowered_scene(tmp_path, "Exact_Translation_Union", "typed_soa"),
            lowered_scene(tmp_path, "Exact_CSG", "canonical_ir")]
    cfgs.append(interpreted(cfgs[0]))
    for k, c in enumerate(cfgs):
        c.width, c.height, c.samples = 48 + 16 * k, 32, 1 + k
    imgs, _ = api.render_batch(cfgs)
    for k, c in enumerate(cfgs):
        assert np.array_equal(imgs[k], api.render(c)), k


def test_learned_instance_below_switch_gives_the_exact_image(tmp_path):
    """The learned march on a lowered scene: a proposer that always answers
    below the switch distance hands every march to the exact distance, so
    the image is bit-identical to the exact kernel's."""
    from test_proposer import _constant_proposer, _plan_image
    c = lowered_scene(tmp_path, "Exact_Translation_Union", "typed_soa")
    c.width, c.height, c.samples = 64, 36, 2
    exact, prof_e, _ = _plan_image(c)
    img, prof_l, learned = _plan_image(c, _constant_proposer(0.01))
    assert np.array_equal(img, exact)
    assert learned["proposals"] > 0 and learned["far_steps"] == 0
    assert prof_l["distance_evals"] == prof_e["distance_evals"]


def _validation_points(c, n=1 << 20):
    """The validator's sample points (Shaders.metal:5106-5124): lowbias32 of
    3 gid + three constants, mapped into the voxel bounds in FP32; points 0-3
    are the origin, the two corners and the centre."""
    import field_parity as fp
    gid = np.arange(n, dtype=np.uint64)
    h = [fp.parity_hash(((gid * 3 + k) & 0xFFFFFFFF).astype(np.uint32)) for k in (0x9E3779B9, 0x243F6A88, 0xB7E15162)]
    unit = np.stack(h, 1).astype(np.float32) * np.float32(1.0 / 4294967295.0)
    lo, hi = np.float32(list(c.voxel_bounds_min)), np.float32(list(c.voxel_bounds_max))
    p = lo + (hi - lo) * unit
    p[0], p[1], p[2], p[3] = 0.0, lo, hi, (lo + hi) * np.float32(0.5)
    return p.astype(np.float64)


def test_typed_soa_validator_passes_the_exact_union_and_counts_a_planted_error(tmp_path):
    """fpt_test_typed_soa on the reference's validator fixture (main.rs:4212-4246,
    Exact_Translation_Union): no failures over 2^20 points.  With one sphere's
    radius grown by 0.01 in the table it fails, and its distance-failure
    count matches the oracle's count at the same points (FP64 distances,
    the same 2e-5 (1 + |d|) test) to within 0.2%."""
    c = lowered_scene(tmp_path, "Exact_Translation_Union", "typed_soa")
    ok, st, msg = api.test_typed_soa(c)
    print("exact union", st)
    assert ok and st["sample_count"] == 1 << 20
    assert st["distance_failures"] == 0 and st["gradient_failures"] == 0
    assert st["max_distance_error"] <= 2e-5 * 8 and st["max_gradient_error"] <= 1e-4

    bad = scenes.clone(c)
    bad.sdf_typed_soa.sphere_radius[0] += 0.01
    ok, st, msg = api.test_typed_soa(bad)
    p = _validation_points(bad)
    ref = interpreted(bad)
    ref.render[6] = bad.render[6] = 0.0
    d_ref, d_tab = bind.map_distance(ref, p, 64), bind.map_distance(bad, p, 64)
    want = int(np.count_nonzero(np.abs(d_ref - d_tab) > 2e-5 * (1 + np.abs(d_ref))))
    print("planted", st, "oracle count", want)
    assert not ok and "typed-SoA validation failed" in msg
    assert want > 1000 and abs(st["distance_failures"] - want) <= 0.002 * want
    assert st["max_distance_error"] == pytest.approx(0.01, rel=1e-3)


def test_typed_soa_validator_flags_a_partial_automatic_table(tmp_path):
    """Automatic typed lowering of a CSG program stops at the first
    intersection and leaves the primitives before it in the table, which the
    renderer then evaluates (scene.rs:1448-1530, 1951); the validator is the
    check that catches it."""
    from fractalpathtracer_amd import scenes as sc
    from test_scene_front import _scene_tree
    path, _ = _scene_tree(tmp_path, "Exact_CSG")
    c, _ = sc.load_scene(path, fpt_root=tmp_path, typed_soa=2)
    assert c.sdf_typed_soa.box_count == 1
    ok, st, msg = api.test_typed_soa(c)
    assert not ok and st["distance_failures"] > 0


@pytest.mark.parametrize("option", ["flat_union", "canonical_ir", "typed_soa"])
def test_voxel_tables_of_a_lowered_scene_are_bit_exact(tmp_path, option):
    """The voxel builder samples the lowered tables, on the GPU and in the
    FP32 oracle alike: page table, material ids and FP16 distances of a
    40^3 sparse 8^3-brick grid are bit-identical, and the voxel render of
    the lowered scene is finite."""
    from fractalpathtracer_amd import ffi
    c = lowered_scene(tmp_path, "Exact_Translation_Union", option)
    c.renderer_backend = ffi.RENDERER_VOXEL
    c.voxel_storage = ffi.VOXEL_STORAGE_SPARSE_BRICKS8
    c.voxel_coverage_mode = 0
    c.voxel_resolution = 40
    c.width, c.height, c.samples = 48, 32, 1
    g = api.voxel_build(c)
    r = bind.voxel_build(c)
    assert g["active_cells"] == r["active_cells"] > 0
    assert np.array_equal(g["page_table"], r["page_table"])
    assert np.array_equal(g["material"], r["material"])
    assert np.array_equal(g["distance"], r["distance"])
    assert np.all(np.isfinite(api.render(c)))
