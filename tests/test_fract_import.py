"""Mandelbulber .fract import for the native Kaleidoscopic IFS field
(fpt_scene_load on a .fract path; src/mandelbulber.rs:226-1041, 1177-1344,
src/scene.rs:1978-2037).  Host-only calls, no GPU.

Pins: the reference's own parser tests (mandelbulber.rs:2564-2830) with
their expected values (decimal commas, IFS masks, legacy field-of-view
migration, global folding slots, world-scale mantissas, target and roll);
the bundled scene ifs-20.fract (kept under tests/golden as a data fixture);
closed forms of the camera (axis swap and world scale, yaw/pitch from the
view direction, the focal length conversion, the distance threshold); and
a geometric check of the roll: the renderer's camera up vector, rotated by
the imported yaw / pitch / roll, is the scene's camera_top projected
orthogonal to the view direction.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from fractalpathtracer_amd import ffi, scenes

GOLDEN = Path(__file__).resolve().parent / "golden"
IFS20 = GOLDEN / "mandelbulber" / "ifs-20.fract"
F32 = lambda x: float(np.float32(x))  # noqa: E731

IFS_SCENE = """
# Mandelbulber settings file
# version 2.33
# only modified parameters
[main_parameters]
camera 0,811466000478122 -0,408023831291236 -1,37676147184911;
camera_top 0,665331264161861 0,700434635203032 -0,258313047941398;
detail_level 2;
formula_1 10;
fov 61,93;
image_height 1080;
image_width 1920;
target 0,810487393068948 -0,405416779822347 -1,3722128396787;
[fractal_1]
IFS_abs_z true;
IFS_direction_2 -0,56 0,61 0,56;
IFS_direction_3 -0,54 0,65 0,54;
IFS_direction_4 0,71 0 -0,71;
IFS_enabled_2 true;
IFS_enabled_3 true;
IFS_enabled_4 true;
IFS_offset 1 1 1;
IFS_rotation -2 8 6;
IFS_rotation_enabled true;
IFS_scale 1,4;
"""


def _load(tmp_path, text, name="s.fract", **opt):
    p = tmp_path / name
    p.write_text(text)
    return scenes.load_scene(p, **opt)


def test_decimal_commas_and_ifs_parameters(tmp_path):
    c, out = _load(tmp_path, IFS_SCENE)
    assert c.sdf_id == ffi.SDF_MANDELBULBER and (c.width, c.height) == (1920, 1080) and out == "s.png"
    sv = list(c.set_values)
    assert sv[0] == 1024 and sv[1] == 250 and sv[2] == 100 and sv[3] == F32(1.4) and sv[39] == 10
    assert sv[4] == 4  # abs mask: z
    assert sv[5] == 0b11100  # IFS_enabled_2..4
    assert sv[6:9] == [1, 1, 1]
    assert sv[9:12] == [F32(math.radians(a)) for a in (-2, 8, 6)]
    for k, d in ((0, (1, 0, 0)), (2, (-0.56, 0.61, 0.56)), (4, (0.71, 0, -0.71))):
        v = np.array(d) / np.linalg.norm(d)
        assert sv[12 + 3 * k:15 + 3 * k] == [F32(x) for x in v]
    assert c.fractal_style_mode == 2 and c.render[1] == 10000 and c.samples == 64


def test_camera_closed_forms_and_roll(tmp_path):
    c, _ = _load(tmp_path, IFS_SCENE)
    cam = np.array([0.811466000478122, -0.408023831291236, -1.37676147184911])
    tgt = np.array([0.810487393068948, -0.405416779822347, -1.3722128396787])
    top = np.array([0.665331264161861, 0.700434635203032, -0.258313047941398])
    sw = lambda v: np.array([v[0], v[2], v[1]])  # noqa: E731
    cw, tw = sw(cam) * 1024, sw(tgt) * 1024
    d = tw - cw
    assert list(c.camera_position) == [F32(x) for x in cw]
    assert c.camera_yaw_pitch[0] == F32(math.atan2(d[0], d[2]))
    assert c.camera_yaw_pitch[1] == F32(math.atan2(d[1], math.hypot(d[0], d[2])))
    assert c.focus_distance == F32(np.linalg.norm(d) * 1024) and c.focus_distance > 5.0  # reference: > 5
    fov = math.radians(61.93)
    assert c.camera_fov == pytest.approx(math.degrees(2 * math.atan(2 * math.tan(fov / 2))), rel=1e-6)
    thr = min(max(np.linalg.norm(d) * (2 * math.tan(fov / 2) / 1080 / 2), 1e-12 * 1024), 1024)
    assert c.render[2] == c.render[3] == F32(thr) and c.render[4] == 50 * 1024
    # roll: rotating the renderer's up (0, 1, 0) by roll, then pitch and yaw, gives the
    # scene's camera_top made orthogonal to the view direction
    yaw, pitch, roll = float(c.camera_yaw_pitch[0]), float(c.camera_yaw_pitch[1]), float(c.camera_roll)
    v = np.array([-math.sin(roll), math.cos(roll), 0.0])
    v = np.array([v[0], v[2] * math.sin(pitch) + v[1] * math.cos(pitch), v[2] * math.cos(pitch) - v[1] * math.sin(pitch)])
    v = np.array([v[0] * math.cos(yaw) + v[2] * math.sin(yaw), v[1], -v[0] * math.sin(yaw) + v[2] * math.cos(yaw)])
    f = d / np.linalg.norm(d)
    up = sw(top) - f * (sw(top) @ f)
    up /= np.linalg.norm(up)
    assert np.allclose(v, up, atol=1e-6) and abs(c.camera_roll) > 1.0  # reference: |roll| > 1


def test_bundled_scene(tmp_path):
    c, out = scenes.load_scene(IFS20)
    ref, _ = _load(tmp_path, IFS_SCENE)
    assert out == "ifs-20.png" and c.set_values[3] == F32(1.4)
    # the bundled file adds the first two directions and non-geometry settings only
    for f in ("camera_position", "camera_yaw_pitch", "render", "world", "sun"):
        assert list(getattr(c, f)) == list(getattr(ref, f)), f
    assert list(c.set_values[3:12]) == list(ref.set_values[3:12])
    assert c.camera_roll == ref.camera_roll and c.camera_fov == ref.camera_fov


def test_legacy_fov_migration(tmp_path):
    legacy = IFS_SCENE.replace("# version 2.33", "# version 2.20").replace("fov 61,93;", "fov 0,5;")
    c, _ = _load(tmp_path, legacy)
    fov = math.degrees(math.atan(0.25)) * 2  # 28.0724869...
    assert fov == pytest.approx(28.072486935853, abs=1e-12)
    assert c.camera_fov == pytest.approx(math.degrees(2 * math.atan(2 * math.tan(math.radians(fov) / 2))), rel=1e-6)
    fish = legacy.replace("detail_level 2;", "detail_level 2;\nperspective_type fish_eye;")
    assert _load(tmp_path, fish)[0].camera_fov == 90.0 and _load(tmp_path, fish)[0].vset_values[107] == 1
    equi = legacy.replace("detail_level 2;", "detail_level 2;\nperspective_type equirectangular;")
    assert _load(tmp_path, equi)[0].camera_fov == 180.0


def test_global_folding_slots_and_direction_normalisation(tmp_path):
    src = IFS_SCENE.replace(
        "[main_parameters]",
        "[main_parameters]\nbox_folding true;\nbox_folding_limit 3;\nbox_folding_value 6;\nspherical_folding true;"
        "\nspherical_folding_outer 2;\nspherical_folding_inner 0,5;").replace(
        "[fractal_1]", "[fractal_1]\nIFS_direction_0 1 2 3;")
    c, _ = _load(tmp_path, src)
    assert list(c.vset_values[101:107]) == [1, 3, 6, 1, 2, 0.5]
    assert c.set_values[12] != 1.0 and c.set_values[12] == F32(1 / math.sqrt(14))


def test_world_scale_preserves_large_coordinates(tmp_path):
    src = IFS_SCENE.replace("camera 0,811466000478122 -0,408023831291236 -1,37676147184911;",
                            "camera 301.6904743885267 -438.2784468930019 -193.1132163166291;").replace(
        "target 0,810487393068948 -0,405416779822347 -1,3722128396787;",
        "target 301.6903192372978 -438.2803842413788 -193.112744501663;")
    c, _ = _load(tmp_path, src)
    rec = [np.float32(x) / np.float32(1024) for x in c.camera_position]
    assert rec == [F32(301.6904743885267), F32(-193.1132163166291), F32(-438.2784468930019)]


def test_options_and_errors(tmp_path):
    c, _ = _load(tmp_path, IFS_SCENE, width=320, height=180, samples=8)
    assert (c.width, c.height, c.samples) == (320, 180, 8)
    full, _ = _load(tmp_path, IFS_SCENE)
    assert c.vset_values[116] == F32(full.vset_values[116] * 1080 / 180)  # threshold scale follows the height
    cases = [
        (IFS_SCENE.replace("formula_1 10;", "formula_1 7;"), "formula ID 7 needs the Mandelbulber2 formula source"),
        (IFS_SCENE.replace("# version 2.33", ""), "missing Mandelbulber version"),
        (IFS_SCENE.replace("# version 2.33", "# version 1.21"), "only Mandelbulber 2.x"),
        (IFS_SCENE + "IFS_scale 2;\n", "duplicate parameter fractal_1.IFS_scale"),
        (IFS_SCENE.replace("detail_level 2;", "detail_level 2"), "missing ';'"),
        (IFS_SCENE.replace("IFS_abs_z true;", "IFS_abs_z yes;"), "invalid fractal_1.IFS_abs_z boolean"),
        (IFS_SCENE.replace("detail_level 2;", "detail_level 2;\nN 5000;"), "N must be 1..4096"),
        (IFS_SCENE.replace("target 0,810487393068948 -0,405416779822347 -1,3722128396787;",
                           "target 0,811466000478122 -0,408023831291236 -1,37676147184911;"),
         "camera and target must differ"),
        (IFS_SCENE.replace("IFS_direction_4 0,71 0 -0,71;", "IFS_direction_4 0 0 0;"), "finite and non-zero"),
        (IFS_SCENE.replace("IFS_scale 1,4;", "IFS_scale 0;"), "IFS_scale must be non-zero"),
        (IFS_SCENE.replace("detail_level 2;", "detail_level 2;\nhybrid_fractal_enable true;"),
         "needs the Mandelbulber2 formula source"),
        (IFS_SCENE.replace("fov 61,93;", "fov 61,93,1;"), "invalid main_parameters.fov"),
    ]
    for text, msg in cases:
        with pytest.raises(ffi.FptError, match=msg):
            _load(tmp_path, text)
    with pytest.raises(ffi.FptError, match="only the direct sdf renderer"):
        _load(tmp_path, IFS_SCENE, renderer_backend=1)
    with pytest.raises(ffi.FptError, match="do not apply to Mandelbulber"):
        _load(tmp_path, IFS_SCENE, flat_union=1)


def test_imported_scene_refusals_on_the_host():
    """Imported Mandelbulber scenes render on the SDF backend's
    one-path-per-lane kernel; the voxel backend, batches and learned
    stepping refuse them before any device work."""
    from fractalpathtracer_amd import api
    c, _ = scenes.load_scene(IFS20)
    c.width, c.height, c.samples = 8, 8, 1
    with pytest.raises(RuntimeError, match="batched rendering"):
        api.Batch([c])
    with pytest.raises(RuntimeError, match="learned stepping"):
        api.proposer_samples(c, 16)
    v = scenes.load_scene(IFS20)[0]
    v.renderer_backend = ffi.RENDERER_VOXEL
    with pytest.raises(RuntimeError, match="SDF backend"):
        api.render(v)
