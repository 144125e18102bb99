"""ctypes binding of the CPU oracle (oracle/liborc.so).

ORACLE — test infrastructure only.  Imported by tests/, by
__graft_entry__.smoke() and by bench.py's CPU-baseline leg, always as the
checker, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liborc.so"
_lib = None


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    args = ["make", "-s", "-C", str(HERE)]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P = C.POINTER
        dp = P(C.c_double)
        L.orc_render.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32, dp]
        L.orc_kifs_sample.argtypes = [C.c_void_p, C.c_int, dp, C.c_size_t, dp]
        L.orc_program_bounds.argtypes = [C.c_void_p, C.c_int, dp, dp, dp]
        L.orc_classify_cell.argtypes = [C.c_void_p, C.c_int, dp, dp, dp]
        L.orc_present.argtypes = [C.c_void_p, C.c_int, P(C.c_float), P(C.c_uint8)]
        L.orc_diagnostic.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_int, C.c_int, P(C.c_uint8)]
        L.orc_render_samples.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_uint32, dp]
        L.orc_trace_primary.argtypes = [C.c_void_p, C.c_int, C.c_int, dp]
        L.orc_environment.argtypes = [C.c_void_p, C.c_int, dp, C.c_size_t, dp]
        L.orc_leaf_refine.argtypes = [C.c_void_p, C.c_int, C.c_uint32, dp, dp]
        L.orc_focus_estimate.argtypes = [C.c_void_p, C.c_int]
        L.orc_focus_estimate.restype = C.c_double
        L.orc_map_distance.argtypes = [C.c_void_p, C.c_int, dp, C.c_size_t, dp]
        L.orc_normal.argtypes = [C.c_void_p, C.c_int, dp, C.c_size_t, dp]
        L.orc_camera_ray.argtypes = [C.c_void_p, C.c_int, P(C.c_uint32), C.c_size_t, dp]
        L.orc_march.argtypes = [C.c_void_p, C.c_int, dp, C.c_size_t, dp]
        L.orc_program_surface.argtypes = [C.c_void_p, C.c_int, dp, C.c_size_t, dp]
        L.orc_leaf_orbit.argtypes = [C.c_uint32, P(C.c_float), C.c_int, dp, dp]
        L.orc_mandelbulb_step.argtypes = [dp, C.c_double]
        L.orc_mandelbox_step.argtypes = [dp, C.c_double, C.c_double, C.c_double]
        L.orc_hash13_bits.argtypes = [C.c_float] * 3
        L.orc_hash13_bits.restype = C.c_uint32
        L.orc_hash13.argtypes = [C.c_float] * 3
        L.orc_hash13.restype = C.c_float
        L.orc_last_counters.argtypes = [P(C.c_uint64)]
        L.orc_random_vector.argtypes = [C.c_int, dp, P(C.c_float), C.c_size_t, dp]
        L.orc_voxel_render_tables.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, P(C.c_uint32),
                                               C.c_size_t, C.c_void_p, C.c_size_t, P(C.c_uint16), P(C.c_uint8),
                                               C.c_size_t, P(C.c_float), C.c_uint32, dp]
        L.orc_scatter.argtypes = [C.c_int, dp, dp]
        _voxel_decls(L)
        _lib = L
    return _lib


def _voxel_decls(L):
    P = C.POINTER
    L.orc_voxel_build.argtypes = [C.c_void_p, C.c_int, P(C.c_uint32), C.c_size_t, C.c_void_p, C.c_size_t,
                                  P(C.c_uint16), C.c_size_t, P(C.c_uint8), C.c_size_t, P(C.c_float),
                                  C.c_size_t, P(C.c_uint64)]
    L.orc_voxel_trace_cells.argtypes = [C.c_void_p, P(C.c_float), C.c_size_t, C.c_uint32, P(C.c_int32),
                                        P(C.c_uint32), P(C.c_float)]
    L.orc_voxel_render.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, P(C.c_double)]
    L.orc_last_voxel_counts.argtypes = [P(C.c_uint64)]
    L.orc_voxel_dda_contract.restype = C.c_uint32
    L.orc_f32_to_f16.argtypes = [C.c_float]
    L.orc_f32_to_f16.restype = C.c_uint16


RECORD_DTYPE = np.dtype([("packed_color", "<u4"), ("packed_properties", "<u4"), ("emission", "<f4")])


def voxel_build(cfg, threads: int = 0) -> dict:
    """FP32 voxel tables: page_table, records or (distance, material, palette)."""
    L = lib()
    info = (C.c_uint64 * 6)()
    if L.orc_voxel_build(C.addressof(cfg), threads, None, 0, None, 0, None, 0, None, 0, None, 0, info) != 0:
        raise RuntimeError("orc_voxel_build failed")
    storage, B, G, pages, cells = (int(info[i]) for i in range(5))
    res = cfg.voxel_resolution
    out = dict(storage=storage, brick_size=B, bricks_per_axis=G, active_bricks=pages, active_cells=cells,
               rejected_bricks=int(info[5]))
    cpb = B ** 3
    pt = np.zeros(G ** 3, np.uint32)
    nrec = res ** 3 if storage == 0 else (pages * cpb if storage == 1 else 0)
    rec = np.zeros(max(nrec, 1), RECORD_DTYPE)
    dist = np.zeros(max(pages * cpb, 1), np.uint16)
    mat = np.zeros(max(pages * cpb, 1), np.uint8)
    pal = np.zeros(256 * 8, np.float32)
    L.orc_voxel_build(C.addressof(cfg), threads, pt.ctypes.data_as(C.POINTER(C.c_uint32)), len(pt),
                      rec.ctypes.data_as(C.c_void_p), len(rec), dist.ctypes.data_as(C.POINTER(C.c_uint16)),
                      len(dist), mat.ctypes.data_as(C.POINTER(C.c_uint8)), len(mat),
                      pal.ctypes.data_as(C.POINTER(C.c_float)), len(pal), info)
    out.update(page_table=pt, records=rec[:nrec], distance=dist[:pages * cpb], material=mat[:pages * cpb],
               palette=pal.reshape(256, 8))
    return out


def voxel_trace_cells(cfg, rays, cap: int = 4096):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    cells = np.zeros((n, cap, 3), np.int32)
    counts = np.zeros(n, np.uint32)
    hits = np.zeros((n, 8), np.float32)
    if lib().orc_voxel_trace_cells(C.addressof(cfg), rays.ctypes.data_as(C.POINTER(C.c_float)), n, cap,
                                   cells.ctypes.data_as(C.POINTER(C.c_int32)),
                                   counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   hits.ctypes.data_as(C.POINTER(C.c_float))) != 0:
        raise RuntimeError("orc_voxel_trace_cells failed")
    return cells, counts, hits


def voxel_render(cfg, threads: int = 0, rows=None) -> np.ndarray:
    y0, y1 = (0, cfg.height) if rows is None else rows
    out = np.zeros((cfg.height, cfg.width, 4), np.float64)
    if lib().orc_voxel_render(C.addressof(cfg), threads, y0, y1, _dp(out)) != 0:
        raise RuntimeError("orc_voxel_render failed")
    return out


def voxel_render_tables(cfg, tables: dict, threads: int = 0, rows=None) -> np.ndarray:
    """renderVoxelPath over tables built elsewhere (a dict like
    api.voxel_build's: page_table, records, distance, material, palette,
    active_bricks)."""
    y0, y1 = (0, cfg.height) if rows is None else rows
    out = np.zeros((cfg.height, cfg.width, 4), np.float64)
    pt = np.ascontiguousarray(tables["page_table"], np.uint32)
    rec = np.ascontiguousarray(tables["records"])
    dist = np.ascontiguousarray(tables["distance"], np.uint16)
    mat = np.ascontiguousarray(tables["material"], np.uint8)
    pal = np.ascontiguousarray(tables["palette"], np.float32).reshape(-1)
    P = C.POINTER
    rc = lib().orc_voxel_render_tables(C.addressof(cfg), threads, y0, y1, pt.ctypes.data_as(P(C.c_uint32)), len(pt),
                                       rec.ctypes.data_as(C.c_void_p), len(rec), dist.ctypes.data_as(P(C.c_uint16)),
                                       mat.ctypes.data_as(P(C.c_uint8)), len(mat), pal.ctypes.data_as(P(C.c_float)),
                                       int(tables["active_bricks"]), _dp(out))
    if rc != 0:
        raise RuntimeError("orc_voxel_render_tables failed")
    return out


def voxel_counts() -> dict:
    """Memory work of the last voxel_render or voxel_trace_cells: cells
    visited, page-table reads (once per brick entered), cell payload reads."""
    out = (C.c_uint64 * 3)()
    lib().orc_last_voxel_counts(out)
    return {"visited_cells": out[0], "page_reads": out[1], "cell_reads": out[2]}


def voxel_dda_contract() -> int:
    return int(lib().orc_voxel_dda_contract())


def f32_to_f16_bits(x: float) -> int:
    return int(lib().orc_f32_to_f16(x))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def render(cfg, precision: int = 64, threads: int = 0, rows=None) -> np.ndarray:
    """Linear float4 accumulation (H, W, 4) in float64; rows=(y0, y1) renders a band."""
    y0, y1 = (0, cfg.height) if rows is None else rows
    out = np.zeros((cfg.height, cfg.width, 4), np.float64)
    rc = lib().orc_render(C.addressof(cfg), precision, threads, y0, y1, _dp(out))
    if rc != 0:
        raise RuntimeError("orc_render failed")
    return out


def counters() -> dict:
    a = (C.c_uint64 * 9)()
    lib().orc_last_counters(a)
    keys = ("distance_evals", "orbit_iterations", "primary_steps", "secondary_steps",
            "shadow_steps", "normal_evals", "paths", "far_steps", "far_orbit_iterations")
    return dict(zip(keys, list(a)))


def leaf_refine(cfg, mode: int, origin, direction, t0, t1, center=(0, 0, 0), half=(0, 0, 0), precision: int = 64):
    """Voxel leaf refinement of one ray interval: (refined?, distance)."""
    ray = np.array([*origin, *direction, t0, t1, *center, *half], np.float64)
    out = np.zeros(1)
    ok = lib().orc_leaf_refine(C.byref(cfg), precision, mode, _dp(ray), _dp(out))
    return bool(ok), float(out[0])


def environment(cfg, dirs, precision: int = 64) -> np.ndarray:
    """environment() of the config for (n, 3) view directions."""
    d = np.ascontiguousarray(np.asarray(dirs, np.float64).reshape(-1, 3))
    out = np.zeros_like(d)
    if lib().orc_environment(C.byref(cfg), precision, _dp(d), len(d), _dp(out)) != 0:
        raise RuntimeError("orc_environment failed")
    return out


def focus_estimate(cfg, precision: int = 64) -> float:
    """estimateFocusDistance of the config's camera (FP64 or the FP32 twin)."""
    return float(lib().orc_focus_estimate(C.byref(cfg), precision))


def trace_primary(cfg, precision: int = 64, threads: int = 0) -> np.ndarray:
    """(H, W, 9): t, px, py, pz, nx, ny, nz, hit, steps."""
    out = np.zeros((cfg.height, cfg.width, 9), np.float64)
    if lib().orc_trace_primary(C.addressof(cfg), precision, threads, _dp(out)) != 0:
        raise RuntimeError("orc_trace_primary failed")
    return out


def render_samples(cfg, samples, rows=None, precision: int = 64, threads: int = 0) -> np.ndarray:
    """Mean of samples range(*samples) per pixel, (H, W, 4) float64 (rows
    outside `rows` stay zero)."""
    y0, y1 = (0, cfg.height) if rows is None else rows
    out = np.zeros((cfg.height, cfg.width, 4), np.float64)
    rc = lib().orc_render_samples(C.addressof(cfg), precision, threads, y0, y1, int(samples[0]), int(samples[1]),
                                  _dp(out))
    if rc != 0:
        raise RuntimeError("orc_render_samples failed")
    return out


def kifs_sample(cfg, points, precision: int = 64) -> np.ndarray:
    """(n, 4) {distance, radius, derivative, +-iterations} of the native KIFS
    field."""
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    out = np.zeros((len(p), 4))
    if lib().orc_kifs_sample(C.addressof(cfg), precision, _dp(p), len(p), _dp(out)) != 0:
        raise RuntimeError("orc_kifs_sample failed")
    return out


def program_bounds(cfg, center, half, precision: int = 64) -> dict:
    c = np.ascontiguousarray(center, np.float64)
    h = np.ascontiguousarray(half, np.float64)
    out = np.zeros(4)
    lib().orc_program_bounds(C.addressof(cfg), precision, _dp(c), _dp(h), _dp(out))
    return {"minimum": out[0], "maximum": out[1], "lipschitz": out[2], "certified": bool(out[3])}


def classify_cell(cfg, center, half, precision: int = 64) -> dict:
    c = np.ascontiguousarray(center, np.float64)
    h = np.ascontiguousarray(half, np.float64)
    out = np.zeros(7)
    lib().orc_classify_cell(C.addressof(cfg), precision, _dp(c), _dp(h), _dp(out))
    return {"interval_min": out[0], "interval_max": out[1], "certified": bool(out[2]), "legacy": bool(out[3]),
            "lipschitz": bool(out[4]), "interval": bool(out[5]), "cell_class": int(out[6])}


def present(cfg, accum, precision: int = 64) -> np.ndarray:
    """RGBA8 (H, W, 4) of a (H, W, 4) float accumulation buffer."""
    a = np.ascontiguousarray(accum, np.float32)
    out = np.zeros((cfg.height, cfg.width, 4), np.uint8)
    if lib().orc_present(C.addressof(cfg), precision, a.ctypes.data_as(C.POINTER(C.c_float)),
                         out.ctypes.data_as(C.POINTER(C.c_uint8))) != 0:
        raise RuntimeError("orc_present failed")
    return out


def diagnostic(cfg, mode: int, bounce: int = 0, max_distance: float = 0.0, precision: int = 64,
               threads: int = 0) -> np.ndarray:
    """Diagnostic RGBA8 (H, W, 4) in PNG row order."""
    out = np.zeros((cfg.height, cfg.width, 4), np.uint8)
    if lib().orc_diagnostic(C.addressof(cfg), int(mode), int(bounce), float(max_distance), precision, threads,
                            out.ctypes.data_as(C.POINTER(C.c_uint8))) != 0:
        raise RuntimeError("orc_diagnostic failed")
    return out


def map_distance(cfg, points, precision: int = 64) -> np.ndarray:
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    out = np.zeros(len(p), np.float64)
    lib().orc_map_distance(C.addressof(cfg), precision, _dp(p), len(p), _dp(out))
    return out


def normal(cfg, points, precision: int = 64) -> np.ndarray:
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    out = np.zeros((len(p), 3), np.float64)
    lib().orc_normal(C.addressof(cfg), precision, _dp(p), len(p), _dp(out))
    return out


def camera_ray(cfg, pixels, precision: int = 64) -> np.ndarray:
    """(n, 6) {u, v, dx, dy, dz, seen} of the un-jittered camera ray of pixels
    (n, 2) {x, y}."""
    xy = np.ascontiguousarray(pixels, np.uint32).reshape(-1, 2)
    out = np.zeros((len(xy), 6), np.float64)
    if lib().orc_camera_ray(C.addressof(cfg), precision, xy.ctypes.data_as(C.POINTER(C.c_uint32)), len(xy),
                            _dp(out)) != 0:
        raise RuntimeError("orc_camera_ray failed")
    return out


def march(cfg, origins, directions, precision: int = 64) -> np.ndarray:
    """(n, 4) {x, y, z, steps}: the scene's march along each ray."""
    o = np.asarray(origins, np.float64).reshape(-1, 3)
    d = np.asarray(directions, np.float64).reshape(-1, 3)
    rays = np.ascontiguousarray(np.concatenate([o, d], axis=1))
    out = np.zeros((len(rays), 4), np.float64)
    if lib().orc_march(C.addressof(cfg), precision, _dp(rays), len(rays), _dp(out)) != 0:
        raise RuntimeError("orc_march failed")
    return out


def program_surface(cfg, points, precision: int = 64) -> np.ndarray:
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    out = np.zeros((len(p), 4), np.float64)
    lib().orc_program_surface(C.addressof(cfg), precision, _dp(p), len(p), _dp(out))
    return out


def leaf_orbit(opcode: int, data, point, precision: int = 64) -> dict:
    d = (C.c_float * 4)(*[float(x) for x in (list(data) + [0.0] * 4)[:4]])
    p = np.ascontiguousarray(point, np.float64)
    out = np.zeros(9, np.float64)
    if lib().orc_leaf_orbit(opcode, d, precision, _dp(p), _dp(out)) != 0:
        raise ValueError("not a fractal leaf opcode")
    return dict(distance=out[0], radius=out[1], derivative=out[2], iterations=int(out[3]),
                escaped=bool(out[4]), z=out[5:8].copy(), trap=out[8])


def mandelbulb_step(z, power: float, derivative: float = 1.0):
    s = np.array([z[0], z[1], z[2], np.sqrt(np.dot(z, z)), derivative], np.float64)
    lib().orc_mandelbulb_step(_dp(s), power)
    return s[:3].copy(), s[4]


def mandelbox_step(z, scale: float, min_r2: float, fixed_r2: float, derivative: float = 1.0):
    s = np.array([z[0], z[1], z[2], np.sqrt(np.dot(z, z)), derivative], np.float64)
    lib().orc_mandelbox_step(_dp(s), scale, min_r2, fixed_r2)
    return s[:3].copy(), s[4]


def hash13_bits(x: float, y: float, z: float) -> int:
    return int(lib().orc_hash13_bits(x, y, z))


def hash13(x: float, y: float, z: float) -> float:
    return float(lib().orc_hash13(x, y, z))


def random_vector(normals, keys, precision: int = 64) -> np.ndarray:
    """randomVector (Shaders.metal:557-568): (n, 3) directions about normals
    (n, 3) with float32 keys (n, 3) = {uv.x, uv.y, seed}."""
    nv = np.ascontiguousarray(normals, np.float64).reshape(-1, 3)
    k = np.ascontiguousarray(keys, np.float32).reshape(-1, 3)
    if len(k) != len(nv):
        raise ValueError("one key per normal")
    out = np.zeros((len(nv), 3), np.float64)
    if lib().orc_random_vector(precision, _dp(nv), k.ctypes.data_as(C.POINTER(C.c_float)), len(nv), _dp(out)):
        raise RuntimeError("orc_random_vector failed")
    return out


def scatter(dr, n, rgb=(1.0, 1.0, 1.0), roughness=1.0, specular=0.0, translucency=0.0, ior=1.5, side=1.0,
            r1=0.5, r2=0.5, diffuse=(0.0, 0.0, 0.0), precision: int = 64) -> dict:
    """One scatter event of renderPath (Shaders.metal:5982-6013)."""
    inp = np.array([*dr, *n, *rgb, roughness, specular, translucency, ior, side, r1, r2, *diffuse], np.float64)
    out = np.zeros(10, np.float64)
    if lib().orc_scatter(precision, _dp(inp), _dp(out)):
        raise RuntimeError("orc_scatter failed")
    return dict(dr=out[0:3].copy(), n=out[3:6].copy(), factor=out[6:9].copy(), side=float(out[9]))
