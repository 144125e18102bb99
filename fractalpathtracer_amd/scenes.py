"""Scene construction: the subset of the reference's scene loader that the
hot path needs, plus the five workloads this library is measured on.

`default_config` restates src/scene.rs:743-776; `apply_scene_json` restates
the camera/render/world/sun/post/voxel/sdf_program parts of
`load_scene_config` (src/scene.rs:1619-1976, program parse 963-1021,
geometry/shading copy 1159-1166).  Presets (.fpt SET values) are not read:
program scenes ignore them.
"""
from __future__ import annotations

import ctypes as C
import json
import math
from pathlib import Path

from . import ffi
from .ffi import FptRenderConfig


def _set(arr, values):
    for i, v in enumerate(values):
        arr[i] = v


def default_config() -> FptRenderConfig:
    """src/scene.rs:743-776."""
    c = FptRenderConfig()
    c.width, c.height, c.samples = 768, 768, 64
    c.sdf_id = ffi.SDF_CORNELL_BOX
    c.sdf_rr_start, c.sdf_rr_min_prob = 3.0, 0.2
    c.mandel_iteration_scale = 1.0
    c.sdf_chunk_samples = 8
    _set(c.camera_position, [0.1, 0.1, -5.0])
    _set(c.camera_yaw_pitch, [0.0, 0.0])
    c.camera_roll, c.camera_fov, c.camera_dof = 0.0, 90.0, 0.01
    _set(c.render, [5.0, 312.0, 0.0005, 0.0005, 1000.0, 0.25, 0.0, 0.0])
    _set(c.world, [0.0, 1.0, 120.0, 30.0, 1.0, 1.0, 0.0])
    _set(c.world_one_color, [1.0] * 3)
    _set(c.sun, [0.0, 120.0, 30.0, 1.0, 0.0])
    _set(c.sun_color, [1.0] * 3)
    _set(c.background_gradient, [1.0] * 6)
    _set(c.post, [0.0, 1.0, 0.0, 1.0, 1.0, 0.0, 0.0])
    _set(c.program_material, [1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 1.5, 0.0])
    _set(c.fractal_style, [0.0, 1.0, 0.5, 1.0, 1.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0])
    c.voxel_resolution = 128
    c.voxel_storage = ffi.VOXEL_STORAGE_SPARSE_BRICKS
    _set(c.voxel_bounds_min, [-4.0] * 3)
    c.voxel_surface_band = 1.0
    _set(c.voxel_bounds_max, [4.0] * 3)
    c.bound_grid_resolution = 32
    c.bound_grid_profile_stride = 4
    c.focus_distance = 0.0
    return c


_OPCODES = {
    "abs": ffi.OP_ABS, "translate": ffi.OP_TRANSLATE, "scale": ffi.OP_SCALE,
    "rotate_x": ffi.OP_ROTATE_X, "rotate_y": ffi.OP_ROTATE_Y, "rotate_z": ffi.OP_ROTATE_Z,
    "repeat": ffi.OP_REPEAT, "sort_desc": ffi.OP_SORT_DESC, "box_fold": ffi.OP_BOX_FOLD,
    "sphere": ffi.OP_SPHERE, "box": ffi.OP_BOX, "plane": ffi.OP_PLANE,
    "mandelbulb": ffi.OP_MANDELBULB, "mandelbox": ffi.OP_MANDELBOX, "menger": ffi.OP_MENGER,
    "union": ffi.OP_UNION, "intersection": ffi.OP_INTERSECTION, "subtract": ffi.OP_SUBTRACT,
    "orbit_add": ffi.OP_ORBIT_ADD,
}
_COMBINE = {"union": 0, "intersection": 1, "intersect": 1, "subtract": 2, "subtraction": 2}


def set_program(c: FptRenderConfig, ops) -> None:
    """Fill sdf_program from a list of (opcode, flags, data4) tuples and copy
    it to the shading program (src/scene.rs:963-990, 1159-1166)."""
    if not 0 < len(ops) <= ffi.MAX_OPS:
        raise ValueError("invalid SDF operation count")
    c.sdf_id = ffi.SDF_PROGRAM
    c.sdf_program_count = len(ops)
    c.sdf_shading_program_count = len(ops)
    for i, (op, flags, data) in enumerate(ops):
        for dst in (c.sdf_program[i], c.sdf_shading_program[i]):
            dst.opcode = op
            dst.flags = flags
            d = list(data) + [0.0] * (4 - len(data))
            _set(dst.data, d[:4])


def load_scene(path, fpt_root=".", **options):
    """Load a JSON scene with the library's C++ front end (fpt_scene_load,
    src/scene.rs:1619-1976): returns (config, output_name).  `options` are
    FptSceneOptions fields (width, height, samples, program_optimization,
    geometry_split, canonical_ir, flat_union, typed_soa, renderer_backend,
    topology_specialization, ...)."""
    lib = ffi.load_library()
    opt = ffi.FptSceneOptions()
    for k, v in options.items():
        setattr(opt, k, int(v))
    cfg = FptRenderConfig()
    name = C.create_string_buffer(512)
    err = ffi.error_buffer()
    ffi.check(lib.fpt_scene_load(str(path).encode(), str(fpt_root).encode(), C.byref(opt), C.byref(cfg), name,
                                 len(name), err, len(err)), err)
    return cfg, name.value.decode()


def set_hdri_path(c: FptRenderConfig, path) -> None:
    """world.hdri (src/scene.rs:873-879): the path as NUL-terminated bytes,
    truncated to 255; HDRI is enabled iff the path is not empty.  The render
    entry points load the file into the LUT when world mode is 2."""
    data = str(path).encode()[: len(c.hdri_path) - 1]
    C.memset(C.addressof(c) + type(c).hdri_path.offset, 0, len(c.hdri_path))
    for i, b in enumerate(data):
        c.hdri_path[i] = b
    c.hdri_enabled = 1 if data else 0


def apply_scene_json(c: FptRenderConfig, scene) -> FptRenderConfig:
    """Apply a scene JSON object (or path) to `c` (src/scene.rs:1619-1976 subset)."""
    base = Path(".")
    if not isinstance(scene, dict):
        base = Path(scene).parent
        scene = json.loads(Path(scene).read_text())
    for k in ("width", "height", "samples"):
        if k in scene:
            setattr(c, k, int(scene[k]))
    cam = scene.get("camera", {})
    if "position" in cam:
        _set(c.camera_position, cam["position"])
    if "yaw_pitch" in cam:
        _set(c.camera_yaw_pitch, cam["yaw_pitch"])
    for k, f in (("roll", "camera_roll"), ("fov", "camera_fov"), ("dof", "camera_dof"),
                 ("focus_distance", "focus_distance")):
        if k in cam:
            setattr(c, f, float(cam[k]))

    def obj(name, target, keys):
        o = scene.get(name)
        if isinstance(o, dict):
            for i, k in enumerate(keys):
                if k in o:
                    target[i] = float(o[k])

    obj("render", c.render, ["bounces", "marching_steps", "normal_epsilon", "min_distance",
                             "max_distance", "adaptive_marching", "error_protection",
                             "random_mode"])
    obj("world", c.world, ["mode", "light_size", "rotation", "elevation", "power", "contrast",
                           "gradient_background"])
    w = scene.get("world", {})
    if "one_color" in w:
        _set(c.world_one_color, w["one_color"])
    if isinstance(w.get("hdri"), str):  # relative to the scene file (src/scene.rs:864-871, 1958-1965)
        h = Path(w["hdri"])
        set_hdri_path(c, h if h.is_absolute() else base / h)
    obj("sun", c.sun, ["enabled", "rotation", "elevation", "power", "softness"])
    if "color" in scene.get("sun", {}):
        _set(c.sun_color, scene["sun"]["color"])
    obj("post", c.post, ["tone_map", "exposure", "brightness", "saturation", "contrast",
                         "chromatic_aberration", "highlights"])
    vox = scene.get("voxel", {})
    if "resolution" in vox:
        c.voxel_resolution = int(vox["resolution"])
    if "bounds_min" in vox:
        _set(c.voxel_bounds_min, vox["bounds_min"])
    if "bounds_max" in vox:
        _set(c.voxel_bounds_max, vox["bounds_max"])
    if "surface_band" in vox:
        c.voxel_surface_band = float(vox["surface_band"])
    prog = scene.get("sdf_program")
    if prog:
        ops = []
        for o in prog["operations"]:
            data = [0.0] * 4
            v = o.get("data", o.get("value"))
            if v is not None:
                for i, x in enumerate(v[:4]):
                    data[i] = float(x)
            for key in ("radius", "angle", "factor"):
                if key in o:
                    data[0] = float(o[key])
            if "orbit_weight" in o:
                data[3] = float(o["orbit_weight"])
            ops.append((_OPCODES[o["op"]], _COMBINE.get(o.get("combine", "union"), 0), data))
        set_program(c, ops)
        mat = prog.get("material", {})
        if "color" in mat:
            _set(c.program_material, mat["color"][:3])
        for i, k in enumerate(["roughness", "specular", "translucency", "ior", "emission"]):
            if k in mat:
                c.program_material[3 + i] = float(mat[k])
        c.material_mode = {"hsv_orbit": 1, "gradient": 2}.get(mat.get("mode"), c.material_mode)
        grad = prog.get("gradient")
        if grad:
            set_gradient(c, [(g["position"], g["color"]) for g in grad])
    return c


def set_gradient(c: FptRenderConfig, stops) -> None:
    c.gradient_count = len(stops)
    for i, (pos, col) in enumerate(stops):
        _set(c.gradient_stops[i], [pos] + list(col))


def _plain_camera(c: FptRenderConfig, w: int, h: int, spp: int) -> None:
    c.width, c.height, c.samples = w, h, spp
    _set(c.camera_position, [0.0, 0.0, -3.0])
    _set(c.camera_yaw_pitch, [0.0, 0.0])
    c.camera_fov, c.camera_dof, c.camera_roll = 60.0, 0.0, 0.0


# Sun direction (0.577, 0.577, -0.577) under rotateCamera (Shaders.metal:460-471):
# pitch = asin(0.577...), yaw = 135 deg.
SUN_ROTATION_DEG = 135.0
SUN_ELEVATION_DEG = math.degrees(math.asin(1.0 / math.sqrt(3.0)))


def _c1_lighting(c: FptRenderConfig) -> None:
    _set(c.sun, [1.0, SUN_ROTATION_DEG, SUN_ELEVATION_DEG, 1.0, 0.0])
    _set(c.sun_color, [1.0, 1.0, 1.0])
    c.world[4] = 0.0  # environment off


def mandelbulb_scene(width=256, height=256, spp=1, bounces=1, power=8.0, iterations=12,
                     bailout=2.0) -> FptRenderConfig:
    """Plain-method workload: Mandelbulb power 8, 12 iterations, bailout 2,
    camera (0,0,-3) fov 60, 256 steps, eps 1e-4, one directional light,
    grey albedo 0.7.  `bounces` counts scatter events after the primary hit,
    so render[0] = bounces + 1."""
    c = default_config()
    _plain_camera(c, width, height, spp)
    _set(c.render, [float(bounces + 1), 256.0, 0.0005, 1e-4, 1000.0, 0.0, 0.0, 0.0])
    _c1_lighting(c)
    set_program(c, [(ffi.OP_MANDELBULB, 0, [power, float(iterations), bailout, 0.0])])
    c.material_mode = 0
    _set(c.program_material, [0.7, 0.7, 0.7, 1.0, 0.0, 0.0, 1.5, 0.0])
    return c


def typed_program_scene(width=1280, height=720, spp=16) -> FptRenderConfig:
    """Typed-program workload: Menger (5 iterations) box-folded and rotated
    30 deg about Y, unioned with a Mandelbox (scale -1.5, 10 iterations,
    minR 0.5, fixedR 1) offset by (2,0,0); 4 bounces with Russian roulette
    after bounce 2; analytic normals; 2-colour gradient over the orbit trap."""
    c = default_config()
    c.width, c.height, c.samples = width, height, spp
    _set(c.camera_position, [1.0, 0.6, -6.0])
    _set(c.camera_yaw_pitch, [0.0, -0.05])
    c.camera_fov, c.camera_dof = 50.0, 0.0
    _set(c.render, [5.0, 256.0, 0.0005, 1e-4, 100.0, 0.0, 0.0, 0.0])
    c.sdf_russian_roulette = 1
    c.sdf_rr_start = 3.0
    c.sdf_normal_mode = 2
    _c1_lighting(c)
    _set(c.world, [3.0, 1.0, 0.0, 0.0, 0.3, 1.0, 0.0])
    _set(c.world_one_color, [0.35, 0.45, 0.6])
    set_program(c, [
        (ffi.OP_TRANSLATE, 0, [2.0, 0.0, 0.0, 0.0]),
        (ffi.OP_MANDELBOX, 0, [-1.5, 10.0, 0.5, 1.0]),
        (ffi.OP_TRANSLATE, 0, [-2.0, 0.0, 0.0, 0.0]),
        (ffi.OP_ROTATE_Y, 0, [math.radians(30.0), 0.0, 0.0, 0.0]),
        (ffi.OP_BOX_FOLD, 0, [1.0, 0.0, 0.0, 0.0]),
        (ffi.OP_MENGER, 0, [5.0, 1.0, 3.0, 0.0]),
    ])
    c.material_mode = 2
    _set(c.program_material, [1.0, 1.0, 1.0, 0.85, 0.2, 0.0, 1.5, 0.0])
    set_gradient(c, [(0.0, [0.85, 0.35, 0.1]), (1.0, [0.1, 0.45, 0.9])])
    return c


def primitive_union_scene(width=1280, height=720, spp=16, lowered=True) -> FptRenderConfig:
    """Lowered-program workload: a translation-only hard union of 28
    primitives, a ground plane y = -1, 15 spheres (radius 0.55-0.8) on a
    5 x 3 grid and 12 boxes (half extents 0.35-0.5) on the 4 x 3 grid between
    them; 3 bounces, analytic normals.  With `lowered` the program is lowered
    to a typed structure-of-arrays table by the library's own pass
    (fpt_scene_lower_typed_soa, scene.rs:1448-1530); without it the
    interpreter evaluates the same program."""
    c = default_config()
    c.width, c.height, c.samples = width, height, spp
    look_at(c, [0.0, 3.2, -6.5], [0.0, -0.2, 2.5])
    c.camera_fov, c.camera_dof = 55.0, 0.0
    _set(c.render, [3.0, 256.0, 0.0005, 1e-4, 100.0, 0.0, 0.0, 0.0])
    c.sdf_normal_mode = 0
    _c1_lighting(c)
    _set(c.world, [3.0, 1.0, 0.0, 0.0, 0.3, 1.0, 0.0])
    _set(c.world_one_color, [0.35, 0.45, 0.6])
    placed = [((0.0, 0.0, 0.0), ffi.OP_PLANE, [0.0, 1.0, 0.0, 1.0])]
    for k in range(15):
        i, j = k % 5, k // 5
        placed.append(((-4.0 + 2.0 * i, -0.2, 2.0 * j), ffi.OP_SPHERE, [0.55 + 0.05 * ((i + j) % 6), 0.0, 0.0, 0.0]))
    for k in range(12):
        i, j = k % 4, k // 4
        h = 0.35 + 0.05 * ((i + 2 * j) % 4)
        placed.append(((-3.0 + 2.0 * i, -1.0 + h, 1.0 + 2.0 * j), ffi.OP_BOX, [h, h, h, 0.0]))
    ops, at = [], (0.0, 0.0, 0.0)
    for pos, op, data in placed:  # translations accumulate through the program
        delta = [p - q for p, q in zip(pos, at)]
        if any(delta):
            ops.append((ffi.OP_TRANSLATE, 0, delta + [0.0]))
        ops.append((op, 0, data))
        at = pos
    set_program(c, ops)
    if lowered and ffi.load_library().fpt_scene_lower_typed_soa(C.byref(c)) != 1:
        raise RuntimeError("the primitive union did not lower to a typed table")
    return c


def learned_step_scene(width=1920, height=1080, spp=64) -> FptRenderConfig:
    """Learned-step workload scene: Mandelbulb power 8, 12 iterations,
    1920x1080, 64 spp, 4 bounces."""
    return mandelbulb_scene(width, height, spp, bounces=4)


def voxel_mandelbox_scene(width=1920, height=1080, spp=4, resolution=1024) -> FptRenderConfig:
    """Voxel workload: built-in Mandelbox with scale 2.0 (set_values[2] =
    -0.8 under Shaders.metal:731), 12 iterations, voxelised over [-7, 7]^3
    (the scale-2 box spans about [-6, 6]^3) on a resolution^3 grid of 8^3
    bricks holding FP16 distance + u8 material id, shell coverage (about 4%
    of cells); 3 bounces."""
    c = default_config()
    c.width, c.height, c.samples = width, height, spp
    c.sdf_id = ffi.SDF_MANDELBOX_FRACTAL
    c.set_values[2] = -0.8
    _set(c.camera_position, [0.0, 0.0, -18.0])
    _set(c.camera_yaw_pitch, [0.4, 0.3])
    c.camera_fov, c.camera_dof = 60.0, 0.0
    _set(c.render, [4.0, 256.0, 0.0005, 1e-4, 100.0, 0.0, 0.0, 0.0])
    _c1_lighting(c)
    _set(c.world, [3.0, 1.0, 0.0, 0.0, 0.4, 1.0, 0.0])
    c.renderer_backend = ffi.RENDERER_VOXEL
    c.voxel_resolution = resolution
    c.voxel_storage = ffi.VOXEL_STORAGE_SPARSE_BRICKS8
    c.voxel_coverage_mode = ffi.VOXEL_COVERAGE_SHELL
    _set(c.voxel_bounds_min, [-7.0] * 3)
    _set(c.voxel_bounds_max, [7.0] * 3)
    c.voxel_surface_band = 1.0
    return c


def _splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def _uniform(seed: int, scene: int, field: int) -> float:
    """Counter-based uniform in [0, 1) of (seed, scene, field)."""
    return _splitmix64((seed << 40) ^ (scene << 16) ^ field) / 2.0 ** 64


def look_at(c: FptRenderConfig, position, target=(0.0, 0.0, 0.0)) -> None:
    """Camera yaw/pitch so that rotateCamera((0,0,1)) points at `target`
    (Shaders.metal:460-471: dir = (cos p sin y, sin p, cos p cos y))."""
    f = [t - p for t, p in zip(target, position)]
    n = math.sqrt(sum(v * v for v in f))
    f = [v / n for v in f]
    _set(c.camera_position, position)
    _set(c.camera_yaw_pitch, [math.atan2(f[0], f[2]), math.asin(f[1])])


def batch_scenes(count=64, seed=0, min_tile=64, max_tile=512, max_spp=16):
    """Batched uneven workload: `count` scenes, family i mod 3 (Mandelbulb
    power U[2,10], Mandelbox scale U[-2.5,2.5], Menger U{3..6} iterations),
    tile width/height U{min_tile..max_tile}, spp U{1..max_spp}, 2 bounces;
    every parameter from a counter-based hash of (seed, scene)."""
    out = []
    for i in range(count):
        u = lambda f: _uniform(seed, i, f)  # noqa: E731
        w = min_tile + int(u(1) * (max_tile - min_tile + 1))
        h = min_tile + int(u(2) * (max_tile - min_tile + 1))
        spp = 1 + int(u(3) * max_spp)
        c = default_config()
        c.width, c.height, c.samples = w, h, spp
        c.camera_fov, c.camera_dof = 60.0, 0.0
        _set(c.render, [3.0, 256.0, 0.0005, 1e-4, 100.0, 0.0, 0.0, 0.0])
        _c1_lighting(c)
        _set(c.world, [3.0, 1.0, 0.0, 0.0, 0.3, 1.0, 0.0])
        _set(c.world_one_color, [0.4, 0.5, 0.65])
        fam = i % 3
        if fam == 0:
            power = 2.0 + 8.0 * u(4)
            set_program(c, [(ffi.OP_MANDELBULB, 0, [power, 12.0, 2.0, 0.0])])
            look_at(c, (0.3, 0.4, -2.8))
        elif fam == 1:
            scale = -2.5 + 5.0 * u(4)
            set_program(c, [(ffi.OP_MANDELBOX, 0, [scale, 10.0, 0.5, 1.0])])
            look_at(c, (1.0, 1.5, -6.5))
        else:
            iters = 3 + int(u(4) * 4)
            set_program(c, [(ffi.OP_ROTATE_Y, 0, [0.5, 0, 0, 0]), (ffi.OP_MENGER, 0, [float(iters), 1.0, 3.0, 0.0])])
            look_at(c, (1.4, 1.1, -3.2))
        c.material_mode = 1
        _set(c.program_material, [0.8, 0.8, 0.8, 0.9, 0.1, 0.0, 1.5, 0.0])
        out.append(c)
    return out


def clone(c: FptRenderConfig) -> FptRenderConfig:
    out = FptRenderConfig()
    C.memmove(C.byref(out), C.byref(c), C.sizeof(FptRenderConfig))
    return out
