"""Host-side calls into libfpt_hip.so, mirroring how the reference's Rust host
drives its bridge (src/main.rs:912-963 `execute_metal_render_internal`,
src/main.rs:3254-3270 field sampling).  Results come back as numpy arrays.
Every call goes through the C ABI; a missing library raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import ffi
from .ffi import FptRenderConfig


@dataclass
class RenderResult:
    linear: np.ndarray            # (H, W, 4) float32 running-mean accumulation
    elapsed_ms: float
    build_ms: float
    profile: dict = field(default_factory=dict)
    voxel: dict = field(default_factory=dict)
    stitch: dict = field(default_factory=dict)


def render_full(cfg: FptRenderConfig, output_png: str | None = None, profile: bool = False) -> RenderResult:
    """fpt_metal_render with every optional output requested."""
    lib = ffi.load_library()
    n = cfg.width * cfg.height * 4
    out = np.zeros(n, np.float32)
    build_ms = C.c_double(0)
    elapsed = C.c_double(0)
    vbytes = C.c_uint64(0)
    vbricks = C.c_uint32(0)
    vcells = C.c_uint64(0)
    vrej = C.c_uint32(0)
    prof = ffi.FptSdfProfileStats()
    cache_status = C.c_uint32(0)
    vstats = ffi.FptStitchValidationStats()
    pstats = ffi.FptStitchPipelineStats()
    err = ffi.error_buffer()
    rc = lib.fpt_metal_render(None, None, None, output_png.encode() if output_png else None, None, 0,
                              C.byref(cfg), C.byref(build_ms), C.byref(elapsed), C.byref(vbytes),
                              C.byref(vbricks), C.byref(vcells), C.byref(vrej), None, C.byref(cache_status),
                              C.byref(vstats), C.byref(pstats),
                              C.byref(prof) if profile else None,
                              out.ctypes.data_as(C.POINTER(C.c_float)), n, err, len(err))
    ffi.check(rc, err)
    stitch = {"cache_status": cache_status.value,
              "validation": {k: getattr(vstats, k) for k, _ in ffi.FptStitchValidationStats._fields_},
              "pipeline": {k: getattr(pstats, k) for k, _ in ffi.FptStitchPipelineStats._fields_}}
    pdict = {}
    if profile:
        pdict = {k: getattr(prof, k) for k, _ in ffi.FptSdfProfileStats._fields_[:13]}
    return RenderResult(out.reshape(cfg.height, cfg.width, 4), elapsed.value, build_ms.value, pdict,
                        dict(memory_bytes=vbytes.value, active_bricks=vbricks.value,
                             active_cells=vcells.value, rejected_bricks=vrej.value), stitch)


def render(cfg: FptRenderConfig) -> np.ndarray:
    return render_full(cfg).linear


def trace_primary(cfg: FptRenderConfig) -> dict:
    lib = ffi.load_library()
    n = cfg.width * cfg.height
    hits = (ffi.FptRayHit * n)()
    err = ffi.error_buffer()
    ffi.check(lib.fpt_hip_trace_primary(C.byref(cfg), hits, n, err, len(err)), err)
    a = np.frombuffer(hits, dtype=np.dtype([("t", "f4"), ("position", "f4", 3), ("normal", "f4", 3),
                                            ("hit", "u4"), ("steps", "u4"), ("_pad0", "u4")]))
    a = a.reshape(cfg.height, cfg.width)
    return {k: np.array(a[k]) for k in ("t", "position", "normal", "hit", "steps")}


def sample_field(cfg: FptRenderConfig, points) -> np.ndarray:
    """fpt_mandelbulber_sample_field: (N, 4) {distance, radius, derivative, iterations}."""
    lib = ffi.load_library()
    p = np.zeros((len(points), 4), np.float32)
    p[:, :3] = np.asarray(points, np.float32).reshape(-1, 3)
    out = (ffi.FptMandelbulberFieldSample * len(p))()
    err = ffi.error_buffer()
    ffi.check(lib.fpt_mandelbulber_sample_field(None, None, 0, C.byref(cfg),
                                                p.ctypes.data_as(C.POINTER(C.c_float)), len(p), out,
                                                err, len(err)), err)
    return np.frombuffer(out, np.float32).reshape(-1, 4).copy()


def program_surface(cfg: FptRenderConfig, points) -> np.ndarray:
    """fpt_hip_program_surface: (N, 4) {distance, gx, gy, gz}."""
    lib = ffi.load_library()
    p = np.zeros((len(points), 4), np.float32)
    p[:, :3] = np.asarray(points, np.float32).reshape(-1, 3)
    out = np.zeros((len(p), 4), np.float32)
    err = ffi.error_buffer()
    ffi.check(lib.fpt_hip_program_surface(C.byref(cfg), p.ctypes.data_as(C.POINTER(C.c_float)), len(p),
                                          out.ctypes.data_as(C.POINTER(C.c_float)), err, len(err)), err)
    return out


RECORD_DTYPE = np.dtype([("packed_color", "<u4"), ("packed_properties", "<u4"), ("emission", "<f4")])


def voxel_build(cfg: FptRenderConfig) -> dict:
    """fpt_hip_voxel_build: the GPU's voxel tables as numpy arrays."""
    lib = ffi.load_library()
    info = ffi.FptVoxelBuildInfo()
    err = ffi.error_buffer()
    ffi.check(lib.fpt_hip_voxel_build(C.byref(cfg), None, 0, None, 0, None, 0, None, 0, None, 0,
                                      C.byref(info), err, len(err)), err)
    B, G, pages = info.brick_size, info.bricks_per_axis, info.active_bricks
    res = cfg.voxel_resolution
    cpb = B ** 3
    pt = np.zeros(max(G ** 3, 1), np.uint32)
    nrec = res ** 3 if info.storage == ffi.VOXEL_STORAGE_DENSE else (
        pages * cpb if info.storage == ffi.VOXEL_STORAGE_SPARSE_BRICKS else 0)
    rec = np.zeros(max(nrec, 1), RECORD_DTYPE)
    dist = np.zeros(max(pages * cpb, 1), np.uint16)
    mat = np.zeros(max(pages * cpb, 1), np.uint8)
    pal = np.zeros(256 * 8, np.float32)
    P = C.POINTER
    ffi.check(lib.fpt_hip_voxel_build(C.byref(cfg), pt.ctypes.data_as(P(C.c_uint32)), len(pt),
                                      rec.ctypes.data_as(C.c_void_p), len(rec),
                                      dist.ctypes.data_as(P(C.c_uint16)), len(dist),
                                      mat.ctypes.data_as(P(C.c_uint8)), len(mat),
                                      pal.ctypes.data_as(P(C.c_float)), len(pal), C.byref(info), err,
                                      len(err)), err)
    return dict(storage=info.storage, brick_size=B, bricks_per_axis=G, active_bricks=pages,
                active_cells=info.active_cells, memory_bytes=info.memory_bytes, build_ms=info.build_ms,
                rejected_bricks=info.rejected_bricks,
                page_table=pt[:G ** 3], records=rec[:nrec], distance=dist[:pages * cpb],
                material=mat[:pages * cpb], palette=pal.reshape(256, 8))


def voxel_trace_cells(cfg: FptRenderConfig, rays, cap: int = 4096):
    """fpt_hip_voxel_trace_cells: (cells (n, cap, 3), counts (n,), hits (n, 8))."""
    lib = ffi.load_library()
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    cells = np.zeros((n, cap, 3), np.int32)
    counts = np.zeros(n, np.uint32)
    hits = np.zeros((n, 8), np.float32)
    err = ffi.error_buffer()
    P = C.POINTER
    ffi.check(lib.fpt_hip_voxel_trace_cells(C.byref(cfg), rays.ctypes.data_as(P(C.c_float)), n, cap,
                                            cells.ctypes.data_as(P(C.c_int32)),
                                            counts.ctypes.data_as(P(C.c_uint32)),
                                            hits.ctypes.data_as(P(C.c_float)), err, len(err)), err)
    return cells, counts, hits


def test_voxel_dda() -> None:
    """fpt_test_voxel_dda (the reference's GPU contract test, ffi.rs:509-520)."""
    lib = ffi.load_library()
    err = ffi.error_buffer(512)
    ffi.check(lib.fpt_test_voxel_dda(None, err, len(err)), err)


def test_typed_soa(cfg: FptRenderConfig) -> tuple[bool, dict, str]:
    """fpt_test_typed_soa (metal_bridge.mm:4116-4203): validate the config's
    typed structure-of-arrays table against its interpreted program at 2^20
    hashed points.  Returns (passed, stats, message); invalid input raises."""
    lib = ffi.load_library()
    st = ffi.FptStitchValidationStats()
    err = ffi.error_buffer()
    rc = lib.fpt_test_typed_soa(None, C.byref(cfg), C.byref(st), err, len(err))
    stats = {f: getattr(st, f) for f, _ in ffi.FptStitchValidationStats._fields_}
    msg = err.value.decode()
    if rc != 0 and st.sample_count == 0:
        raise RuntimeError(msg)
    return rc == 0, stats, msg


def hydrate_hdri(cfg: FptRenderConfig) -> FptRenderConfig:
    """A copy of `cfg` with its HDRI LUT loaded from hdri_path, as every
    render entry point does (fpt_hip_hydrate_hdri; host only)."""
    lib = ffi.load_library()
    out = FptRenderConfig.from_buffer_copy(cfg)
    err = ffi.error_buffer()
    ffi.check(lib.fpt_hip_hydrate_hdri(C.byref(out), err, len(err)), err)
    return out


def topology_source(cfg: FptRenderConfig) -> str:
    """Generated topology-specialised source of cfg's program."""
    lib = ffi.load_library()
    need = C.c_size_t(0)
    err = ffi.error_buffer()
    ffi.check(lib.fpt_hip_topology_source(C.byref(cfg), None, 0, C.byref(need), err, len(err)), err)
    buf = C.create_string_buffer(need.value)
    ffi.check(lib.fpt_hip_topology_source(C.byref(cfg), buf, need.value, None, err, len(err)), err)
    return buf.value.decode()


def proposer_samples(cfg: FptRenderConfig, count: int, seed: int = 0, box_half: float = 1.5,
                     band: float = 0.05, max_attempts: int = 1 << 16) -> tuple[np.ndarray, int]:
    """Training samples (count, 4) = (x, y, z, exact d) and the number of
    band points that found no candidate (fpt_hip_proposer_samples)."""
    lib = ffi.load_library()
    out = np.zeros((int(count), 4), np.float32)
    fails = C.c_uint32(0)
    err = ffi.error_buffer()
    ffi.check(lib.fpt_hip_proposer_samples(C.byref(cfg), box_half, band, int(count), int(seed), int(max_attempts),
                                           out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(fails), err,
                                           len(err)), err)
    return out, int(fails.value)


def proposer_eval(params, points, box_half: float = 1.5) -> np.ndarray:
    """Proposer network on the GPU at (n, 3) points (fpt_hip_proposer_eval)."""
    lib = ffi.load_library()
    pts = np.asarray(points, np.float32).reshape(-1, 3)
    p4 = np.zeros((pts.shape[0], 4), np.float32)
    p4[:, :3] = pts
    w = np.ascontiguousarray(params, np.float32).ravel()
    out = np.zeros(pts.shape[0], np.float32)
    err = ffi.error_buffer()
    ffi.check(lib.fpt_hip_proposer_eval(w.ctypes.data_as(C.POINTER(C.c_float)), w.size, box_half,
                                        p4.ctypes.data_as(C.POINTER(C.c_float)), pts.shape[0],
                                        out.ctypes.data_as(C.POINTER(C.c_float)), err, len(err)), err)
    return out


def diagnostic(cfg: FptRenderConfig, mode, max_distance: float = 0.0, bounce: int = 0,
               output_png: str | None = None) -> tuple[np.ndarray, float]:
    """Diagnostic image (H, W, 4) uint8 in PNG row order and the GPU time in
    ms (fpt_hip_diagnostic_render; with output_png, fpt_metal_diagnostic_render
    writes the PNG instead and the image is read back from it)."""
    lib = ffi.load_library()
    m = ffi.DIAGNOSTIC_MODES[mode] if isinstance(mode, str) else int(mode)
    dg = ffi.FptDiagnosticConfig(m, int(bounce), float(max_distance), 0.0)
    ms = C.c_double(0)
    err = ffi.error_buffer()
    if output_png:
        ffi.check(lib.fpt_metal_diagnostic_render(None, output_png.encode(), None, 0, C.byref(cfg), C.byref(dg),
                                                  C.byref(ms), err, len(err)), err)
        return read_png(output_png), ms.value
    out = np.zeros((cfg.height, cfg.width, 4), np.uint8)
    ffi.check(lib.fpt_hip_diagnostic_render(C.byref(cfg), C.byref(dg), out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                            out.size, C.byref(ms), err, len(err)), err)
    return out, ms.value


def read_png(path: str) -> np.ndarray:
    """Decode an 8-bit RGBA, non-interlaced PNG into (H, W, 4) uint8 (all five
    row filters; the library's writer uses filter 0)."""
    import struct
    import zlib
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG")
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            if depth != 8 or ctype != 6 or body[12] != 0:
                raise ValueError(f"{path}: only 8-bit RGBA non-interlaced PNGs are read")
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    out = np.zeros((h, 4 * w), np.uint8)
    prev = np.zeros(4 * w, np.int32)
    for y in range(h):
        f, line = int(raw[y, 0]), raw[y, 1:].astype(np.int32)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 255
        else:
            cur = np.zeros(4 * w, np.int32)
            for x in range(4 * w):
                a = int(cur[x - 4]) if x >= 4 else 0
                b = int(prev[x])
                c = int(prev[x - 4]) if x >= 4 else 0
                if f == 1:
                    pred = a
                elif f == 3:
                    pred = (a + b) // 2
                else:
                    p = a + b - c
                    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                    pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                cur[x] = (int(line[x]) + pred) & 255
        out[y] = cur
        prev = cur.astype(np.int32)
    return out.reshape(h, w, 4)


def device_name() -> str:
    lib = ffi.load_library()
    buf = C.create_string_buffer(256)
    lib.fpt_metal_device_name(buf, len(buf))
    return buf.value.decode()


class DeviceBuffer:
    """A float32 device buffer owned by the library's HIP runtime."""

    def __init__(self, count: int):
        self.lib = ffi.load_library()
        self.count = int(count)
        self.ptr = C.c_void_p()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_device_alloc(self.count * 4, C.byref(self.ptr), err, len(err)), err)

    def fill(self, value: float, stream: "Stream | None" = None) -> None:
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_fill_f32(self.ptr, value, self.count, stream.handle if stream else None, err,
                                            len(err)), err)

    def numpy(self) -> np.ndarray:
        out = np.empty(self.count, np.float32)
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_copy(out.ctypes.data_as(C.c_void_p), self.ptr, self.count * 4, 0, err,
                                        len(err)), err)
        return out

    @property
    def address(self) -> int:
        return self.ptr.value or 0

    def close(self) -> None:
        if self.ptr:
            self.lib.fpt_hip_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_copy(dst: DeviceBuffer, src: DeviceBuffer) -> None:
    """Device-to-device copy of src into dst by the library's copy kernel."""
    n = min(dst.count, src.count) * 4
    err = ffi.error_buffer()
    ffi.check(dst.lib.fpt_hip_copy(dst.ptr, src.ptr, n - n % 16, 2, err, len(err)), err)


class Stream:
    """A non-blocking HIP stream of the library's runtime."""

    def __init__(self):
        self.lib = ffi.load_library()
        self.handle = C.c_void_p()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_stream_create(C.byref(self.handle), err, len(err)), err)

    @property
    def address(self) -> int:
        return self.handle.value or 0

    def synchronize(self) -> None:
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_synchronize(self.handle, err, len(err)), err)

    def close(self) -> None:
        if self.handle:
            self.lib.fpt_hip_stream_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def time_renders(target, buffer: DeviceBuffer, stream: Stream, steps: int) -> float:
    """HIP-event time (ms) of `steps` renders of a Plan or Batch on `stream`."""
    lib = ffi.load_library()
    ms = C.c_double(0)
    err = ffi.error_buffer()
    plan = target.handle if isinstance(target, Plan) else None
    batch = target.handle if isinstance(target, Batch) else None
    ffi.check(lib.fpt_hip_time_renders(plan, batch, buffer.ptr, stream.handle, steps, C.byref(ms), err, len(err)),
              err)
    return ms.value


class Plan:
    """Device-resident render plan (fpt_hip_plan_*): upload once, render many."""

    def __init__(self, cfg: FptRenderConfig):
        self.lib = ffi.load_library()
        self.cfg = cfg
        self.handle = C.c_void_p()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_plan_create(C.byref(cfg), C.byref(self.handle), err, len(err)), err)

    def tune(self, kernel: int = 1, workgroups: int = 0, event_batch: int = 0) -> None:
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_plan_tune(self.handle, kernel, workgroups, event_batch, err, len(err)), err)

    def tune_stash(self, refill: int = 6, sample_groups: int = 8) -> None:
        """Non-marching lanes before a wave refills them from its LDS path
        stash, and sample runs per pixel (fpt_hip_plan_tune_stash)."""
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_plan_tune_stash(self.handle, refill, sample_groups, err, len(err)), err)

    def set_tile_order(self, mode: int) -> None:
        """1 longest-first (default), 0 row-major (fpt_hip_set_tile_order)."""
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_set_tile_order(self.handle, None, mode, err, len(err)), err)

    def autotune(self, device_ptr: int, stream_ptr: int = 0, rounds: int = 3) -> ffi.FptTuneResult:
        res = ffi.FptTuneResult()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_autotune(self.handle, None, C.c_void_p(device_ptr), C.c_void_p(stream_ptr),
                                            rounds, C.byref(res), err, len(err)), err)
        return res

    def render_into(self, device_ptr: int, stream_ptr: int = 0) -> None:
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_plan_render(self.handle, C.c_void_p(device_ptr), C.c_void_p(stream_ptr),
                                               err, len(err)), err)

    def set_proposer(self, params, box_half: float = 1.5, switch_distance: float = 0.02,
                     resume_distance: float = 0.1, step_scale: float = 0.5) -> None:
        """Attach a learned step proposer (params: FPT_PROPOSER_PARAM_COUNT
        floats), or detach it with params=None."""
        err = ffi.error_buffer()
        if params is None:
            ffi.check(self.lib.fpt_hip_plan_set_proposer(self.handle, None, 0, None, err, len(err)), err)
            return
        w = np.ascontiguousarray(params, np.float32).ravel()
        st = ffi.FptProposerSettings(box_half, switch_distance, resume_distance, step_scale)
        ffi.check(self.lib.fpt_hip_plan_set_proposer(self.handle, w.ctypes.data_as(C.POINTER(C.c_float)), w.size,
                                                     C.byref(st), err, len(err)), err)

    def specialize(self, enable: bool = True) -> dict:
        """Compile (or fetch from the cache), validate and attach the
        topology-specialised kernel (fpt_hip_plan_specialize); the stats
        are returned, and a failed validation raises FptError."""
        st = ffi.FptSpecializationStats()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_plan_specialize(self.handle, int(enable), C.byref(st), err, len(err)), err)
        return {k: getattr(st, k) for k, _ in ffi.FptSpecializationStats._fields_}

    def specialized_surface(self, points) -> np.ndarray:
        """fpt_hip_plan_specialized_surface: (N, 4) {distance, gx, gy, gz} of
        the attached topology-specialised code."""
        p = np.zeros((len(points), 4), np.float32)
        p[:, :3] = np.asarray(points, np.float32).reshape(-1, 3)
        out = np.zeros((len(p), 4), np.float32)
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_plan_specialized_surface(self.handle, p.ctypes.data_as(C.POINTER(C.c_float)),
                                                            len(p), out.ctypes.data_as(C.POINTER(C.c_float)), err,
                                                            len(err)), err)
        return out

    def profile(self, buffer: "DeviceBuffer") -> tuple[dict, dict]:
        """Counting pass of the full frame: (profile counters, learned-step
        counters)."""
        prof = ffi.FptSdfProfileStats()
        lst = ffi.FptProposerStats()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_plan_profile(self.handle, buffer.ptr, C.byref(prof), C.byref(lst), err,
                                                len(err)), err)
        return ({k: getattr(prof, k) for k, _ in ffi.FptSdfProfileStats._fields_[:13]},
                {k: getattr(lst, k) for k, _ in ffi.FptProposerStats._fields_})

    def lane_stats(self, buffer: "DeviceBuffer") -> dict:
        """Lane occupancy of the persistent kernel over the full frame
        (fpt_hip_plan_lane_stats)."""
        st = ffi.FptLaneStats()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_plan_lane_stats(self.handle, buffer.ptr, C.byref(st), err, len(err)), err)
        return {k: getattr(st, k) for k, _ in ffi.FptLaneStats._fields_[:-1]}

    def voxel_stats(self, buffer: "DeviceBuffer") -> dict:
        """Counting pass of a voxel plan's full frame: cells visited, page-table
        entries and cell payloads read by every traversal."""
        st = ffi.FptVoxelTraceStats()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_plan_voxel_stats(self.handle, buffer.ptr, C.byref(st), err, len(err)), err)
        return {k: getattr(st, k) for k, _ in ffi.FptVoxelTraceStats._fields_}

    def close(self) -> None:
        if self.handle:
            self.lib.fpt_hip_plan_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched multi-scene render (fpt_hip_batch_*)."""

    def __init__(self, configs):
        self.lib = ffi.load_library()
        self.configs = list(configs)
        arr = (FptRenderConfig * len(self.configs))(*self.configs)
        self._arr = arr
        self.handle = C.c_void_p()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_batch_create(arr, len(self.configs), C.byref(self.handle), err, len(err)), err)
        self.pixels = int(self.lib.fpt_hip_batch_pixels(self.handle))
        self.offsets = np.cumsum([0] + [c.width * c.height for c in self.configs])

    def tune(self, workgroups: int = 0, event_batch: int = 0, use_graph: bool = True) -> None:
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_batch_tune(self.handle, workgroups, event_batch, int(use_graph), err, len(err)),
                  err)

    def set_tile_order(self, mode: int) -> None:
        """1 longest-first (default), 0 row-major (fpt_hip_set_tile_order)."""
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_set_tile_order(None, self.handle, mode, err, len(err)), err)

    def autotune(self, device_ptr: int, stream_ptr: int = 0, rounds: int = 3) -> ffi.FptTuneResult:
        res = ffi.FptTuneResult()
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_autotune(None, self.handle, C.c_void_p(device_ptr), C.c_void_p(stream_ptr),
                                            rounds, C.byref(res), err, len(err)), err)
        return res

    def render_into(self, device_ptr: int, stream_ptr: int = 0) -> None:
        err = ffi.error_buffer()
        ffi.check(self.lib.fpt_hip_batch_render(self.handle, C.c_void_p(device_ptr), C.c_void_p(stream_ptr), err,
                                                len(err)), err)

    def split(self, flat: np.ndarray):
        """Per-scene (H, W, 4) views of a flat (pixels, 4) result."""
        flat = flat.reshape(-1, 4)
        return [flat[self.offsets[i]:self.offsets[i + 1]].reshape(c.height, c.width, 4)
                for i, c in enumerate(self.configs)]

    def close(self) -> None:
        if self.handle:
            self.lib.fpt_hip_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_batch(configs):
    """fpt_hip_render_batch: list of (H, W, 4) images and the elapsed ms."""
    lib = ffi.load_library()
    configs = list(configs)
    arr = (FptRenderConfig * len(configs))(*configs)
    n = sum(c.width * c.height for c in configs)
    out = np.zeros(n * 4, np.float32)
    ms = C.c_double(0)
    err = ffi.error_buffer()
    ffi.check(lib.fpt_hip_render_batch(arr, len(configs), out.ctypes.data_as(C.POINTER(C.c_float)), out.size,
                                       C.byref(ms), err, len(err)), err)
    flat = out.reshape(-1, 4)
    res, o = [], 0
    for c in configs:
        res.append(flat[o:o + c.width * c.height].reshape(c.height, c.width, 4))
        o += c.width * c.height
    return res, ms.value
