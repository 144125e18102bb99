"""Field-parity harness for imported Mandelbulber scenes: the library's FP32
field (fpt_mandelbulber_sample_field) against the oracle's double-precision
field, restating the reference's `mandel-parity` command
(src/main.rs:3057-3428).

Test infrastructure: it compares the product with the oracle and is imported
only by tests.  Run it by hand on a GPU as

    python tests/field_parity.py scene.fract [--samples N] [--report path.json]

Method (main.rs:3207-3361):
- sample 0..7 are the origin, the scene's camera, its target, their midpoint
  and four cube corners; the rest are hashed into [-2, 2]^3 Mandelbulber
  units with the lowbias32 integer hash (main.rs:3057-3063);
- each point is scaled by the world scale, its y and z swapped and rounded
  to FP32 for the GPU; the CPU side evaluates that same FP32 point, mapped
  back, in double precision;
- distances are compared in scene units (the GPU's divided by the world
  scale): a sample fails above 5e-5 * (1 + |d_cpu|); the field is
  geometry-qualified when no sample is non-finite, at most 2e-4 of them fail
  and the 99th-percentile error (nearest-rank, main.rs:3065-3072) is at most
  5e-5; orbit-qualified when, in addition, at most 2e-4 of the iteration
  counts differ.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ALLOWED_OUTLIER_FRACTION = 2.0e-4
HASH_OFFSETS = (0x9E3779B9, 0x243F6A88, 0xB7E15162)  # main.rs:3228


def parity_hash(v) -> np.ndarray:
    """lowbias32 (main.rs:3057-3063) on uint32 arrays, wrapping multiplies."""
    v = np.asarray(v, np.uint32).copy()
    v ^= v >> np.uint32(16)
    v *= np.uint32(0x7FEB352D)
    v ^= v >> np.uint32(15)
    v *= np.uint32(0x846CA68B)
    return v ^ (v >> np.uint32(16))


def percentile(values, fraction: float) -> float:
    """Nearest-rank percentile of main.rs:3065-3072."""
    v = np.sort(np.asarray(values, np.float64))
    if len(v) == 0:
        return 0.0
    return float(v[int(round((len(v) - 1) * fraction))])


def fract_vec3(path, key: str, default):
    """`key x y z;` from a .fract file's [main_parameters], decimal commas
    accepted (mandelbulber.rs:659-660 defaults)."""
    section = None
    for raw in Path(path).read_text().splitlines():
        line = raw.strip()
        if line.startswith("[") and line.endswith("]"):
            section = line[1:-1]
            continue
        if section != "main_parameters" or not line.startswith(key + " "):
            continue
        words = line[len(key):].strip().rstrip(";").split()
        return [float(w.replace(",", ".")) for w in words[:3]]
    return list(default)


def parity_points(camera, target, world_scale: float, count: int):
    """(gpu (count, 3) float32 world points, cpu (count, 3) float64 Mandelbulber
    points) of main.rs:3207-3256."""
    if not 16 <= count <= 1_048_576:
        raise ValueError("mandel-parity samples must be 16..1048576")
    cam, tgt = np.asarray(camera, np.float64), np.asarray(target, np.float64)
    src = np.empty((count, 3))
    src[:8] = [[0, 0, 0], cam, tgt, (cam + tgt) * 0.5, [-1, -1, -1], [1, 1, 1], [-1, 1, -1], [1, -1, 1]]
    idx = np.arange(8, count, dtype=np.uint32)
    for axis in range(3):
        h = parity_hash(idx * np.uint32(3) + np.uint32(axis) + np.uint32(HASH_OFFSETS[axis]))
        src[8:, axis] = h.astype(np.float64) * (4.0 / 4294967295.0) - 2.0
    gpu = (src[:, [0, 2, 1]] * world_scale).astype(np.float32)
    cpu = gpu.astype(np.float64)[:, [0, 2, 1]] / world_scale
    return gpu, cpu


def parity_report(gpu_samples, cpu_samples, world_scale: float) -> dict:
    """The comparison and qualification of main.rs:3287-3361.  gpu_samples is
    (n, 4) {distance, radius, derivative, iterations} with the distance in
    world units; cpu_samples the same in scene units (iterations signed or
    not)."""
    g = np.asarray(gpu_samples, np.float64)
    c = np.asarray(cpu_samples, np.float64)
    n = len(g)
    finite = np.isfinite(g[:, :3]).all(axis=1)
    g, c = g[finite], c[finite]
    d_err = np.abs(c[:, 0] - g[:, 0] / world_scale)
    r_err = np.abs(c[:, 1] - g[:, 1])
    k_err = np.abs(c[:, 2] - g[:, 2])
    failures = int(np.sum(d_err > 5.0e-5 * (1.0 + np.abs(c[:, 0]))))
    same = np.abs(c[:, 3]) == np.abs(g[:, 3])
    mismatches = int(np.sum(~same))
    rel = lambda e, ref: float(np.max(e / (1.0 + np.abs(ref)), initial=0.0))  # noqa: E731
    p99 = percentile(d_err, 0.99)
    geometry = bool(finite.all() and failures / n <= ALLOWED_OUTLIER_FRACTION and p99 <= 5.0e-5)
    return {
        "sample_count": n,
        "metal_fpt_world_scale": world_scale,
        "distance_tolerance": "5e-5 * (1 + abs(cpu_distance))",
        "mean_absolute_distance_error": float(d_err.sum() / max(len(d_err), 1)),
        "p99_absolute_distance_error": p99,
        "max_absolute_distance_error": float(np.max(d_err, initial=0.0)),
        "max_absolute_radius_error": float(np.max(r_err, initial=0.0)),
        "max_absolute_derivative_error": float(np.max(k_err, initial=0.0)),
        "max_same_iteration_radius_relative_error": rel(r_err[same], c[same, 1]),
        "max_same_iteration_derivative_relative_error": rel(k_err[same], c[same, 2]),
        "distance_failures": failures,
        "distance_failure_fraction": failures / n,
        "iteration_mismatches": mismatches,
        "iteration_mismatch_fraction": mismatches / n,
        "non_finite_samples": int(n - finite.sum()),
        "allowed_outlier_fraction": ALLOWED_OUTLIER_FRACTION,
        "geometry_qualified": geometry,
        "orbit_qualified": bool(geometry and mismatches / n <= ALLOWED_OUTLIER_FRACTION),
        "qualified": geometry,
    }


def cpu_field(cfg, cpu_points, world_scale: float, precision: int = 64) -> np.ndarray:
    """The oracle's KIFS field at the CPU points, distances in scene units.
    The oracle takes renderer (world) coordinates; scaling by a power of two
    and swapping axes are exact, so the points it sees are the GPU's."""
    from oracle import bind
    world = np.asarray(cpu_points, np.float64)[:, [0, 2, 1]] * world_scale
    out = bind.kifs_sample(cfg, world, precision)
    out[:, 0] /= world_scale
    return out


def run(scene_path, samples: int = 65536, report_path=None) -> dict:
    """mandel-parity on an imported .fract scene: GPU field vs the oracle."""
    from fractalpathtracer_amd import api, scenes
    cfg, _ = scenes.load_scene(scene_path)
    world_scale = float(cfg.set_values[0])
    camera = fract_vec3(scene_path, "camera", (0.0, 0.0, -5.0))
    target = fract_vec3(scene_path, "target", (0.0, 0.0, 0.0))
    gpu_pts, cpu_pts = parity_points(camera, target, world_scale, samples)
    t0 = time.perf_counter()
    gpu = api.sample_field(cfg, gpu_pts)
    t1 = time.perf_counter()
    cpu = cpu_field(cfg, cpu_pts, world_scale)
    report = parity_report(gpu, cpu, world_scale)
    t2 = time.perf_counter()
    report = {"scene": str(scene_path), "backend": "mandelbulber-handwritten-kaleidoscopic-ifs-hip-f32",
              "reference": "cpu-f64", "sample_bounds_mandelbulber": [[-2.0] * 3, [2.0] * 3],
              "gpu_batch_elapsed_ms": (t1 - t0) * 1e3, "cpu_reference_and_comparison_elapsed_ms": (t2 - t1) * 1e3,
              **report}
    if report_path is not None:
        Path(report_path).parent.mkdir(parents=True, exist_ok=True)
        Path(report_path).write_text(json.dumps(report, indent=2) + "\n")
    return report


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("scene")
    ap.add_argument("--samples", type=int, default=65536)
    ap.add_argument("--report", default="reports/mandel/field-parity.json")
    a = ap.parse_args(argv)
    r = run(a.scene, a.samples, a.report)
    print(f"field parity: {r['sample_count']} samples, p99 {r['p99_absolute_distance_error']:.3e}, "
          f"max {r['max_absolute_distance_error']:.3e}, failures {r['distance_failures']}, "
          f"iteration mismatches {r['iteration_mismatches']}, qualified {r['qualified']}; wrote {a.report}",
          file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    raise SystemExit(main())
