"""Throughput benchmark: path samples/s of the HIP path tracer on MI355X.

Workload (learned-step scene at full size): Mandelbulb power 8, 12 DE
iterations, bailout 2, 1920x1080, 64 spp, 4 bounces, one directional light,
grey albedo 0.7.  One step = one full frame (64 spp) rendered from the
device-resident plan into a device accumulation buffer; inputs are resident
in HBM before timing starts.

Also reported on the same JSON line:
  roofline      algorithmic FLOPs of the dominant kernel per launch (from an
                untimed counting pass of the same frame) / its average launch
                time measured with HIP events on the launch stream;
  cpu_baseline  the FP64 CPU oracle on the host cores, on a bounded band of
                rows of the same frame.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "path samples/s at 1080p: learned-step vs exact-DE march at equal image RMSE"

# Algorithmic operation counts (FLOPs; each transcendental counts 1):
#  Mandelbulb orbit iteration in the reference's triplex form
#  (orbit.rs:766-778 + add_c + |z| + bailout test): 22 arithmetic + 7
#  transcendentals (asin, atan2, pow, 2 cos, 2 sin).
FLOPS_PER_ORBIT_ITERATION = 29
#  Per distance evaluation outside the orbit loop: Logarithmic finalizer
#  (log + 3), error-protection mix (3), march step (abs, scale, 3-vector
#  advance = 8, LOD/threshold/max tests 3).
FLOPS_PER_DISTANCE_EVAL = 17
#  Analytic normal = forward-mode gradient: 4 lanes of value+3 tangents,
#  counted as 4 distance evaluations of the same orbit.
NORMAL_DE_FACTOR = 4

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 peak


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=4)
    return ap.parse_args()


def cpu_baseline(cfg, rows: int) -> dict:
    """FP64 oracle on the host cores over `rows` rows at the frame centre."""
    from oracle import bind
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    y0 = cfg.height // 2 - rows // 2
    t = time.perf_counter()
    bind.render(cfg, 64, threads=threads, rows=(y0, y0 + rows))
    dt = time.perf_counter() - t
    samples = rows * cfg.width * cfg.samples
    return {"value": samples / dt, "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": f"FP64 C++ oracle, rows {y0}-{y0 + rows - 1} of the {cfg.width}x{cfg.height} frame at "
                      f"{cfg.samples} spp ({samples} path samples, {dt:.1f} s)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU split of this path "
              "is a design note, not a build target", file=sys.stderr)
        sys.exit(2)
    from fractalpathtracer_amd import api, scenes

    name = api.device_name()
    assert "gfx" in name, f"bench.py needs a HIP GPU ({name})"
    cfg = scenes.learned_step_scene(args.width, args.height, args.spp)

    # Untimed counting pass of the same frame: work counters for the roofline.
    prof = api.render_full(cfg, profile=True).profile

    # Timed region: HIP events on the launch stream around `steps` frames,
    # inputs (config) resident in HBM, output in a device buffer.
    stream = api.Stream()
    accum = api.DeviceBuffer(cfg.width * cfg.height * 4)
    plan = api.Plan(cfg)
    if args.warmup:
        api.time_renders(plan, accum, stream, args.warmup)
    total_ms = api.time_renders(plan, accum, stream, args.steps)
    plan.close()
    ms_per_step = total_ms / args.steps
    samples_per_step = cfg.width * cfg.height * cfg.samples
    value = samples_per_step / (ms_per_step / 1e3)

    # One launch renders `chunk` samples of the frame (8 at 64 spp).
    chunk = 8 if cfg.samples > 16 else cfg.samples
    launches = (cfg.samples + chunk - 1) // chunk
    launch_ms = ms_per_step / launches
    iters = prof.get("march_orbit_iterations", 0)
    de = prof.get("distance_evals", 0)
    normals = prof.get("normal_evals", 0)
    avg_iter = iters / max(de, 1)
    flops_frame = (iters * FLOPS_PER_ORBIT_ITERATION + de * FLOPS_PER_DISTANCE_EVAL +
                   normals * NORMAL_DE_FACTOR * (avg_iter * FLOPS_PER_ORBIT_ITERATION + FLOPS_PER_DISTANCE_EVAL))
    flops_launch = flops_frame / launches
    achieved = flops_launch / (launch_ms / 1e3) / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "path samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"exact-DE march, Mandelbulb p8/12it/bailout2, {cfg.width}x{cfg.height}, "
                               f"{cfg.samples} spp, 4 bounces", "method": "exact-DE",
                   "width": cfg.width, "height": cfg.height, "spp": cfg.samples, "bounces": 4,
                   "learned_step": None},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "sdf_accumulate_kernel", "launch_ms": launch_ms,
                     "flops_per_launch": flops_launch,
                     "counts_per_frame": {"orbit_iterations": iters, "distance_evals": de,
                                          "normal_evals": normals}},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(cfg, args.cpu_rows)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
