"""Throughput benchmark: path samples/s of the HIP path tracer on MI355X,
exact-distance march vs learned-step march at equal image RMSE.

Workload (learned-step scene at full size): Mandelbulb power 8, 12 DE
iterations, bailout 2, 1920x1080, 64 spp, 4 bounces, one directional light,
grey albedo 0.7.  One step = one full frame (64 spp) rendered from the
device-resident plan into a device accumulation buffer; inputs (config,
proposer weights) are resident in HBM before timing starts.

Both march methods are timed on the same plan:
  exact    every march step on the exact distance estimate;
  learned  far-field steps (|q| > 0.02 inside the [-1.5, 1.5]^3 training box)
           at 0.5 x the learned proposal q (fractalpathtracer_amd/assets/
           proposer_mandelbulb_p8.npz), exact near the surface.
Image quality of each method is measured on the linear accumulation buffer
over a band of rows against the FP64 CPU oracle twice:
  same samples    the oracle's samples 0..spp-1 (identical RNG keys): parity,
                  how closely the GPU path follows the reference's paths;
  converged       the oracle's samples spp..spp+ref_spp-1 (independent RNG
                  keys, ref_spp = 4 x spp): RMSE against an independent
                  higher-spp estimate, i.e. Monte-Carlo error + bias.
`value` is the throughput of the faster method whose converged RMSE is
within 10% of the exact method's (equal image RMSE); both methods' figures
are on the line.

Also on the JSON line:
  roofline      algorithmic FLOPs of the dominant kernel (the persistent
                path-tracing kernel of the reported method) per launch, from
                an untimed counting pass of the same frame, / its average
                launch time measured with HIP events on the launch stream;
  cpu_baseline  the FP64 CPU oracle on the host cores, timed on the converged
                reference band (same scene, ref_spp samples per pixel).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "path samples/s at 1080p: learned-step vs exact-DE march at equal image RMSE"

# Algorithmic operation counts.  Arithmetic FLOPs (add/sub/mul/min/max/cmp 1,
# FMA 2, div 1, sqrt 1) are the roofline's numerator; transcendentals are
# tallied apart, against the quarter-rate special-function peak.
#  Mandelbulb orbit iteration in the reference's triplex form
#  (orbit.rs:766-778 + add_c + |z| + bailout test): 22 arithmetic + 7
#  transcendentals (asin, atan2, pow, 2 cos, 2 sin).
FLOPS_PER_ORBIT_ITERATION = 22
SPECIAL_PER_ORBIT_ITERATION = 7
#  Per distance evaluation outside the orbit loop: Logarithmic finalizer
#  (3 + log), error-protection mix (3), march step (abs, scale, 3-vector
#  advance = 8, LOD/threshold/max tests 3).
FLOPS_PER_DISTANCE_EVAL = 16
SPECIAL_PER_DISTANCE_EVAL = 1
#  Analytic normal = forward-mode gradient: 4 lanes of value+3 tangents,
#  counted as 4 distance evaluations of the same orbit.
NORMAL_DE_FACTOR = 4
#  Proposer network per point: 2 x (128*64 + 3*64*64 + 64) multiply-adds.
FLOPS_PER_PROPOSAL = 2 * (128 * 64 + 3 * 64 * 64 + 64)

# HBM traffic of one exact-march frame launch, measured with rocprofv3 PMC
# passes (profiles/sdf_persistent_mandelbulb_1080p_64spp.md): FETCH_SIZE and
# WRITE_SIZE KiB per dispatch, scaled by what the same counters read for a
# 1 GiB copy kernel (x2.000 and x1.000).
TRAFFIC_FETCH_KIB, TRAFFIC_WRITE_KIB = 784982.625, 18028032.65625
TRAFFIC_FETCH_SCALE, TRAFFIC_WRITE_SCALE = 2.0, 1.0
TRAFFIC_WORKLOAD = (1920, 1080, 64)

# Typed-program scene: orbit iterations are a mix of Mandelbox (box fold 12,
# |z|^2 5, sphere fold 6, scale + offset 6, derivative 2 = 31 FLOP) and
# Menger (abs 3, sort 6, scale + offset 6, cut 2 = 17 FLOP) iterations that
# the counting pass does not separate; each is charged the smaller figure,
# so the reported FLOP rate is a lower bound.  Per distance evaluation
# outside the leaves: translate 3 + rotate 6 + box fold 12 + translate 3 +
# union 1 + the march step's 8 + 3.
TYPED_FLOPS_PER_ORBIT_ITERATION = 17
TYPED_FLOPS_PER_DISTANCE_EVAL = 36
# Voxel DDA: algorithmic bytes = 4 B per page-table entry read (once per
# brick entered) + 1 B per sparse-bricks8 material id read.
VOXEL_PAGE_ENTRY_BYTES, VOXEL_CELL_BYTES = 4, 1
# HBM traffic of one voxel frame launch (1920x1080, 4 spp), FETCH_SIZE and
# WRITE_SIZE KiB per dispatch (profiles/voxel_persistent_mandelbox_1024_1080p_4spp.md),
# scaled like the default line's.
VOXEL_TRAFFIC_FETCH_KIB, VOXEL_TRAFFIC_WRITE_KIB = 95512.375, 76903.0
VOXEL_TRAFFIC_WORKLOAD = (1920, 1080, 4, 1024)
# Batched workload: arithmetic FLOPs per orbit iteration per leaf family.
# Mandelbulb of a general power in the triplex form (as above, 22 + 7
# transcendentals), Mandelbox (box fold 12, |z|^2 5, sphere fold 6, scale +
# offset 6, derivative 2 = 31), Menger (abs 3, sort 6, scale + offset 6, cut
# 2 = 17); per distance evaluation outside the orbit 16 + a logarithm for the
# Mandelbulb, 17 for the Mandelbox finaliser, plus 6 for the Menger scenes'
# rotation.
BATCH_FLOPS_PER_ORBIT_ITERATION = (22, 31, 17)
BATCH_FLOPS_PER_DISTANCE_EVAL = (16, 17, 23)
HBM_PEAK_GBS = 8000.0      # MI355X published HBM3E bandwidth
FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak
# Transcendentals issue at a quarter of the unpacked FP32 rate: 256 CUs x
# 16 per clock x 2.4 GHz.
SPECIAL_PEAK_TOPS = FP32_PEAK_TFLOPS / 16.0
BF16_PEAK_TFLOPS = 2500.0  # MI355X published dense BF16 matrix peak
KERNEL = "sdf_persistent_kernel"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=8)
    ap.add_argument("--ref-spp-factor", type=int, default=4)
    ap.add_argument("--workload", choices=("learned", "typed", "voxel", "fract", "lowered", "batch"),
                    default="learned",
                    help="learned: the learned-step scene above (the default line); typed: the typed-program "
                         "scene (1280x720, 16 spp); voxel: the 1024^3 sparse-brick Mandelbox (1920x1080, 4 spp); "
                         "fract: the bundled Mandelbulber scene ifs-20.fract (Kaleidoscopic IFS, its own "
                         "1920x1080, 64 spp); lowered: a 28-primitive translated union lowered to a typed "
                         "structure-of-arrays table (1280x720, 16 spp), timed beside the interpreter; batch: 64 "
                         "uneven scenes (Mandelbulb / Mandelbox / Menger, 64^2-512^2 tiles, 1-16 spp, 2 bounces) "
                         "in one graph-replayed persistent launch with the autotuned shape")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy kernel of MIB MiB (the known-size reference for "
                         "HBM byte counters) and exit")
    return ap.parse_args()


def oracle_band(cfg, rows: int, ref_factor: int):
    """FP64 oracle over `rows` rows at the frame centre: the same-sample band
    and the converged band (timed: the CPU baseline)."""
    from oracle import bind
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    y0 = cfg.height // 2 - rows // 2
    y1 = y0 + rows
    same = bind.render(cfg, 64, threads=threads, rows=(y0, y1))[y0:y1]
    s0, s1 = cfg.samples, cfg.samples * (1 + ref_factor)
    t = time.perf_counter()
    conv = bind.render_samples(cfg, (s0, s1), rows=(y0, y1), threads=threads)[y0:y1]
    dt = time.perf_counter() - t
    counts = bind.counters()
    samples = rows * cfg.width * (s1 - s0)
    base = {"value": samples / dt, "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": f"FP64 C++ oracle (OpenMP), rows {y0}-{y1 - 1} of the {cfg.width}x{cfg.height} frame, "
                      f"samples {s0}-{s1 - 1} per pixel ({samples} path samples, {dt:.1f} s)",
            "cpu_model": cpu_model(), "single_core": single_core_rate(cfg, y0)}
    return base, same, conv, (y0, y1), counts


def frame_flops(prof: dict) -> float:
    return counted_flops(prof, FLOPS_PER_ORBIT_ITERATION, FLOPS_PER_DISTANCE_EVAL)


def frame_special_ops(prof: dict) -> float:
    return counted_flops(prof, SPECIAL_PER_ORBIT_ITERATION, SPECIAL_PER_DISTANCE_EVAL)


def tune_dict(t) -> dict:
    """The autotuner's choice (fpt_hip_autotune) as JSON; event batch 0 is
    the launch's default (64 paths for a frame in sample runs, 48 without)."""
    return {"workgroups": t.workgroups, "event_batch": t.event_batch or "launch default",
            "tile_order": "longest-first" if t.tile_order else "row-major", "best_ms": t.best_ms,
            "default_ms": t.default_ms, "candidates": t.candidates, "rounds": t.rounds,
            **({"sample_runs": t.sample_groups} if t.sample_groups else {})}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def single_core_rate(cfg, y0: int) -> dict:
    """The FP64 oracle on one host core: one row, 16 samples per pixel."""
    from oracle import bind
    t = time.perf_counter()
    bind.render_samples(cfg, (cfg.samples, cfg.samples + 16), rows=(y0, y0 + 1), threads=1)
    dt = time.perf_counter() - t
    n = cfg.width * 16
    return {"value": n / dt, "cores": 1, "sample": f"row {y0}, samples {cfg.samples}-{cfg.samples + 15} "
                                                    f"({n} path samples, {dt:.1f} s)"}


def run_method(api, plan, cfg, stream, accum, steps, warmup):
    """Timed frames, then an untimed counting pass and the final image."""
    if warmup:
        api.time_renders(plan, accum, stream, warmup)
    total_ms = api.time_renders(plan, accum, stream, steps)
    img = accum.numpy().reshape(cfg.height, cfg.width, 4)
    prof, learned = plan.profile(accum)
    return total_ms / steps, img, prof, learned


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU split of this path "
              "is a design note, not a build target", file=sys.stderr)
        sys.exit(2)
    from fractalpathtracer_amd import api, proposer, scenes

    name = api.device_name()
    assert "gfx" in name, f"bench.py needs a HIP GPU ({name})"
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src, dst = api.DeviceBuffer(n), api.DeviceBuffer(n)
        src.fill(1.0)
        api.device_copy(dst, src)
        print(json.dumps({"copy_kernel_bytes_read": n * 4, "copy_kernel_bytes_written": n * 4}), flush=True)
        return
    if args.workload == "batch":
        run_batch(args, api, scenes)
        return
    if args.workload != "learned":
        run_workload(args, api, scenes)
        return
    cfg = scenes.learned_step_scene(args.width, args.height, args.spp)
    samples_per_step = cfg.width * cfg.height * cfg.samples
    launches = 1   # the persistent kernel renders all samples of the frame in one launch

    stream = api.Stream()
    accum = api.DeviceBuffer(cfg.width * cfg.height * 4)
    plan = api.Plan(cfg)
    tune = plan.autotune(accum.address, stream.address, rounds=2)
    methods = {}
    ms, img, prof, _ = run_method(api, plan, cfg, stream, accum, args.steps, args.warmup)
    methods["exact"] = {"ms_per_step": ms, "value": samples_per_step / (ms / 1e3), "img": img, "prof": prof}
    if not args.no_learned:
        params, meta = proposer.load(proposer.asset_path("mandelbulb_p8"))
        plan.set_proposer(params, box_half=meta.get("box_half", proposer.BOX_HALF),
                          switch_distance=proposer.SWITCH_DISTANCE, resume_distance=proposer.RESUME_DISTANCE,
                          step_scale=proposer.STEP_SCALE)
        # The exact march's event batch counts lanes and parked paths; the
        # learned instance keeps no parked paths and takes its own default.
        plan.tune(1, tune.workgroups, 0)
        ms, img, prof, learned = run_method(api, plan, cfg, stream, accum, args.steps, args.warmup)
        plan.set_proposer(None)
        methods["learned"] = {"ms_per_step": ms, "value": samples_per_step / (ms / 1e3), "img": img,
                              "prof": prof, "learned": learned, "train": meta}
    plan.close()

    base = None
    gate = None
    band_counts = None
    chosen = "exact"
    if not args.no_cpu_baseline:
        base, same, conv, (y0, y1), band_counts = oracle_band(cfg, args.cpu_rows, args.ref_spp_factor)
        for m in methods.values():
            band = m["img"][y0:y1, :, :3]
            m["rmse_same"] = float(np.sqrt(np.mean((band - same[..., :3]) ** 2)))
            m["rmse"] = float(np.sqrt(np.mean((band - conv[..., :3]) ** 2)))
            m["mean_bias"] = float(np.mean(band) - np.mean(conv[..., :3]))
        gate = 1.10 * methods["exact"]["rmse"]
        ok = [k for k, m in methods.items() if m["rmse"] <= gate]
        chosen = max(ok, key=lambda k: methods[k]["value"])

    m = methods[chosen]
    traffic = None
    if chosen == "exact" and (cfg.width, cfg.height, cfg.samples) == TRAFFIC_WORKLOAD:
        traffic = 1024 * (TRAFFIC_FETCH_KIB * TRAFFIC_FETCH_SCALE + TRAFFIC_WRITE_KIB * TRAFFIC_WRITE_SCALE)
    launch_ms = m["ms_per_step"] / launches
    flops_launch = frame_flops(m["prof"]) / launches
    special_launch = frame_special_ops(m["prof"]) / launches
    achieved = flops_launch / (launch_ms / 1e3) / 1e12
    summary = {}
    for k, mm in methods.items():
        s = {"value": mm["value"], "ms_per_step": mm["ms_per_step"],
             "rmse_vs_converged_fp64": mm.get("rmse"), "rmse_vs_fp64_same_samples": mm.get("rmse_same"),
             "mean_minus_converged": mm.get("mean_bias"),
             "distance_evals": mm["prof"]["distance_evals"],
             "orbit_iterations": mm["prof"]["march_orbit_iterations"]}
        if k == "learned":
            L = mm["learned"]
            # Useful work: one network evaluation per proposal a lane used;
            # issued work: every wave-wide call computes 64 points.
            secs = mm["ms_per_step"] / 1e3
            s.update(L)
            s["mfma_tflops"] = L["proposals"] * FLOPS_PER_PROPOSAL / secs / 1e12
            s["mfma_frac_of_bf16_peak"] = s["mfma_tflops"] / BF16_PEAK_TFLOPS
            s["mfma_tflops_issued"] = L["proposer_calls"] * 64 * FLOPS_PER_PROPOSAL / secs / 1e12
            s["proposals_per_call"] = L["proposals"] / max(L["proposer_calls"], 1)
            if band_counts:
                # Even a free proposer only removes the exact far-field
                # steps (|d| > 0.02): the bound on its speed-up.
                share = band_counts["far_orbit_iterations"] / max(band_counts["orbit_iterations"], 1)
                s["learned_ceiling"] = {"far_field_orbit_iteration_share": share, "speedup_bound": 1.0 / (1.0 - share),
                                        "source": "FP64 oracle work counters on the converged band"}
            s["speedup_vs_exact"] = mm["value"] / methods["exact"]["value"]
            s["training"] = {k2: mm["train"].get(k2) for k2 in
                             ("samples", "steps", "sample_seconds", "train_seconds", "mae_far")}
        summary[k] = s
    line = {
        "metric": METRIC,
        "value": m["value"],
        "unit": "path samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Mandelbulb p8/12it/bailout2, {cfg.width}x{cfg.height}, {cfg.samples} spp, "
                               f"4 bounces", "method": chosen, "rmse_gate": gate,
                   "width": cfg.width, "height": cfg.height, "spp": cfg.samples, "bounces": 4,
                   "autotuned": tune_dict(tune), "methods": summary},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": KERNEL, "launches_per_step": launches, "launch_ms": launch_ms,
                     "flops_per_launch": flops_launch,
                     "special_ops": {"per_launch": special_launch,
                                     "achieved_tops": special_launch / (launch_ms / 1e3) / 1e12,
                                     "peak_tops": SPECIAL_PEAK_TOPS,
                                     "frac": special_launch / (launch_ms / 1e3) / 1e12 / SPECIAL_PEAK_TOPS}},
    }
    if base is not None:
        line["cpu_baseline"] = base
    print(json.dumps(line), flush=True)


def one_path_launches(cfg) -> int:
    """Launches of the one-path-per-lane kernel per frame: the library's
    accumulation policy (one launch up to 16 spp in auto mode, else chunks of
    sdf_chunk_samples, default 8)."""
    from fractalpathtracer_amd import ffi
    spp = min(max(cfg.samples, 1), 512)
    mode = cfg.sdf_accumulation_mode
    if mode == ffi.SDF_ACCUMULATION_PER_SAMPLE:
        return spp
    if mode == ffi.SDF_ACCUMULATION_BATCH or (mode == ffi.SDF_ACCUMULATION_AUTO and spp <= 16):
        return 1
    chunk = min(max(cfg.sdf_chunk_samples or 8, 1), 64)
    return -(-spp // chunk)


def counted_flops(prof: dict, per_iteration: float, per_distance_eval: float) -> float:
    """Algorithmic FLOPs of a counting pass: orbit iterations, distance
    evaluations, and each analytic normal as NORMAL_DE_FACTOR evaluations of
    the mean orbit."""
    iters = prof.get("march_orbit_iterations", 0)
    de = prof.get("distance_evals", 0)
    avg_iter = iters / max(de, 1)
    return (iters * per_iteration + de * per_distance_eval +
            prof.get("normal_evals", 0) * NORMAL_DE_FACTOR * (avg_iter * per_iteration + per_distance_eval))


def run_batch(args, api, scenes):
    """C3: 64 uneven scenes in one persistent launch, captured in a HIP graph
    and replayed, at the autotuned launch shape."""
    cfgs = scenes.batch_scenes(64, seed=0)
    samples_per_step = sum(c.width * c.height * c.samples for c in cfgs)
    stream = api.Stream()
    batch = api.Batch(cfgs)
    accum = api.DeviceBuffer(batch.pixels * 4)
    tune = batch.autotune(accum.address, stream.address, rounds=2)
    if args.warmup:
        api.time_renders(batch, accum, stream, args.warmup)
    ms = api.time_renders(batch, accum, stream, args.steps) / args.steps
    # One launch per scene (each scene its own graph-free plan), for scale.
    solo_ms, flops, counts = 0.0, 0.0, {"distance_evals": 0, "march_orbit_iterations": 0, "normal_evals": 0}
    for i, c in enumerate(cfgs):
        buf = api.DeviceBuffer(c.width * c.height * 4)
        with api.Plan(c) as plan:
            solo_ms += api.time_renders(plan, buf, stream, 1)
            prof, _ = plan.profile(buf)
        fam = i % 3
        flops += counted_flops(prof, BATCH_FLOPS_PER_ORBIT_ITERATION[fam], BATCH_FLOPS_PER_DISTANCE_EVAL[fam])
        for k in counts:
            counts[k] += prof[k]
    achieved = flops / (ms / 1e3) / 1e12
    line = {"metric": "path samples/s, batch workload", "value": samples_per_step / (ms / 1e3),
            "unit": "path samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": "batch: 64 scenes (family i mod 3: Mandelbulb power U[2,10], Mandelbox scale "
                                   "U[-2.5,2.5], Menger U{3..6} iterations), tiles U{64..512}^2, spp U{1..16}, "
                                   "2 bounces, seed 0; one graph-replayed persistent launch",
                       "scenes": len(cfgs), "pixels": batch.pixels, "samples_per_step": samples_per_step,
                       "autotuned": tune_dict(tune),
                       "one_launch_per_scene_ms": solo_ms, "speedup_vs_one_launch_per_scene": solo_ms / ms,
                       "counts": counts},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": KERNEL,
                         "launches_per_step": 1, "launch_ms": ms, "flops_per_launch": flops}}
    if not args.no_cpu_baseline:
        from oracle import bind
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        t = time.perf_counter()
        n = 0
        for c in cfgs:
            y0 = c.height // 2 - 1
            bind.render(c, 64, threads=threads, rows=(y0, y0 + 2))
            n += 2 * c.width * c.samples
        dt = time.perf_counter() - t
        line["cpu_baseline"] = {"value": n / dt, "unit": "path samples/s", "cores": threads, "kind": "port",
                                "sample": f"FP64 C++ oracle (OpenMP), the two centre rows of each of the 64 "
                                          f"scenes at its own spp ({n} path samples, {dt:.1f} s)",
                                "cpu_model": cpu_model()}
    batch.close()
    print(json.dumps(line), flush=True)


def run_workload(args, api, scenes):
    """One JSON line for the typed-program, voxel, imported-scene or
    lowered-program workload, with the same contract keys as the default
    line (its own roofline)."""
    from fractalpathtracer_amd import ffi
    if args.workload == "typed":
        cfg = scenes.typed_program_scene()
        desc = "typed program: Menger(5) box-folded and rotated 30 deg about Y, union Mandelbox(-1.5, 10) " \
               "offset (2,0,0), 1280x720, 16 spp, 4 bounces, Russian roulette after bounce 2"
    elif args.workload == "lowered":
        cfg = scenes.primitive_union_scene()
        desc = "lowered program: translated union of a ground plane, 15 spheres and 12 boxes as a typed " \
               "structure-of-arrays table, 1280x720, 16 spp, 3 bounces"
    elif args.workload == "fract":
        cfg, _ = scenes.load_scene(ROOT / "tests" / "golden" / "mandelbulber" / "ifs-20.fract")
        desc = "Mandelbulber ifs-20.fract: Kaleidoscopic IFS (250 iterations), Mandelbulber camera and march, " \
               f"{cfg.width}x{cfg.height}, {cfg.samples} spp, {int(cfg.render[0])} bounces"
    else:
        cfg = scenes.voxel_mandelbox_scene()
        desc = "voxel: Mandelbox(2.0, 12) on a 1024^3 grid of sparse 8^3 bricks (shell coverage), " \
               "exact DDA, 1920x1080, 4 spp, 3 bounces"
    samples_per_step = cfg.width * cfg.height * cfg.samples
    stream = api.Stream()
    accum = api.DeviceBuffer(cfg.width * cfg.height * 4)
    t0 = time.perf_counter()
    plan = api.Plan(cfg)
    build_s = time.perf_counter() - t0
    tune = None
    if args.workload in ("typed", "lowered"):  # the persistent-kernel workloads
        tune = plan.autotune(accum.address, stream.address, rounds=2)
    if args.warmup:
        api.time_renders(plan, accum, stream, args.warmup)
    ms = api.time_renders(plan, accum, stream, args.steps) / args.steps
    line = {"metric": f"path samples/s, {args.workload} workload", "value": samples_per_step / (ms / 1e3),
            "unit": "path samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": desc, "width": cfg.width, "height": cfg.height, "spp": cfg.samples,
                       "plan_create_s": build_s}}
    if tune is not None:
        line["config"]["autotuned"] = tune_dict(tune)
    if args.workload == "fract":
        prof, _ = plan.profile(accum)
        # per KIFS iteration: 3 abs folds, 3 plane reflections (dot 5, test 1,
        # reflect 7 each), rotation about an offset (3 + 9 + 3 + 3), scaling
        # (6), derivative (1), |z| (6) = 73 FLOP with every plane enabled;
        # per distance evaluation the finaliser (4), threshold (9) and step (8)
        flops = prof["march_orbit_iterations"] * 73 + prof["distance_evals"] * 21
        achieved = flops / (ms / 1e3) / 1e12
        line["roofline"] = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                            "kernel": KERNEL, "launches_per_step": 1,
                            "launch_ms": ms, "flops_per_launch": flops}
        line["config"]["counts"] = {k: prof[k] for k in ("distance_evals", "march_orbit_iterations",
                                                         "normal_evals", "primary_steps", "shadow_steps")}
        # The same frame on the one-path-per-lane kernel (plan kernel 0).
        with api.Plan(cfg) as oplan:
            oplan.tune(kernel=0)
            if args.warmup:
                api.time_renders(oplan, accum, stream, args.warmup)
            oms = api.time_renders(oplan, accum, stream, args.steps) / args.steps
        line["config"]["one_path_kernel"] = {"ms_per_step": oms, "value": samples_per_step / (oms / 1e3),
                                             "launches_per_step": one_path_launches(cfg)}
    elif args.workload == "lowered":
        prof, _ = plan.profile(accum)
        t = cfg.sdf_typed_soa
        # per distance evaluation of the table: sphere 11 FLOP (offset 3, dot
        # 5, sqrt, radius, min), box 23 (offset 3, abs 3, extent 3, max/min 3,
        # clamp 3, length 6, add, min), plane 19 (normalise 9, offset 3,
        # dot 5, add, min)
        per_eval = 11 * t.sphere_count + 23 * t.box_count + 19 * t.plane_count
        flops = prof["distance_evals"] * per_eval
        achieved = flops / (ms / 1e3) / 1e12
        line["roofline"] = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": KERNEL,
                            "flops_per_launch": flops, "flops_per_distance_eval": per_eval, "lower_bound": True}
        interp = scenes.primitive_union_scene(lowered=False)
        with api.Plan(interp) as iplan:
            # the table plan's autotuned launch shape, for a like-for-like comparison
            iplan.tune(1, tune.workgroups, tune.event_batch)
            iplan.set_tile_order(tune.tile_order)
            if args.warmup:
                api.time_renders(iplan, accum, stream, args.warmup)
            ims = api.time_renders(iplan, accum, stream, args.steps) / args.steps
            iprof, _ = iplan.profile(accum)
            # the program's topology compiled by hipRTC into straight-line
            # distance and surface code (validated at 2^20 points first)
            t0 = time.perf_counter()
            spec = iplan.specialize()
            spec_s = time.perf_counter() - t0
            if args.warmup:
                api.time_renders(iplan, accum, stream, args.warmup)
            sms = api.time_renders(iplan, accum, stream, args.steps) / args.steps
        line["config"]["interpreted"] = {"ms_per_step": ims, "value": samples_per_step / (ims / 1e3),
                                         "distance_evals": iprof["distance_evals"]}
        line["config"]["speedup_vs_interpreter"] = ims / ms
        line["config"]["specialized"] = {"ms_per_step": sms, "value": samples_per_step / (sms / 1e3),
                                         "specialize_s": spec_s, "cache_status": spec.get("cache_status"),
                                         "typed_table_speedup_over_specialized": sms / ms}
        line["config"]["counts"] = {k: prof[k] for k in ("distance_evals", "normal_evals", "primary_steps",
                                                         "shadow_steps")}
    elif args.workload == "typed":
        prof, _ = plan.profile(accum)
        flops = (prof["march_orbit_iterations"] * TYPED_FLOPS_PER_ORBIT_ITERATION +
                 prof["distance_evals"] * TYPED_FLOPS_PER_DISTANCE_EVAL)
        achieved = flops / (ms / 1e3) / 1e12
        line["roofline"] = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": KERNEL,
                            "flops_per_launch": flops, "lower_bound": True}
        line["config"]["counts"] = {k: prof[k] for k in ("distance_evals", "march_orbit_iterations",
                                                         "normal_evals", "primary_steps", "shadow_steps")}
        # The same program's topology compiled by hipRTC into straight-line
        # distance and surface code (validated at 2^20 points first), timed
        # at the autotuned launch shape; the interpreter stays the line's value.
        with api.Plan(cfg) as splan:
            t0 = time.perf_counter()
            spec = splan.specialize()
            spec_s = time.perf_counter() - t0
            splan.tune(1, tune.workgroups, tune.event_batch)
            splan.set_tile_order(tune.tile_order)
            splan.tune_stash(sample_groups=tune.sample_groups)
            if args.warmup:
                api.time_renders(splan, accum, stream, args.warmup)
            sms = api.time_renders(splan, accum, stream, args.steps) / args.steps
        line["config"]["specialized"] = {"ms_per_step": sms, "value": samples_per_step / (sms / 1e3),
                                         "specialize_s": spec_s, "cache_status": spec.get("cache_status"),
                                         "speedup_over_interpreter": ms / sms}
        if not args.no_cpu_baseline:
            base, _, _, _, _ = oracle_band(cfg, args.cpu_rows, 1)
            line["cpu_baseline"] = base
    else:
        st = plan.voxel_stats(accum)
        assert cfg.voxel_storage == ffi.VOXEL_STORAGE_SPARSE_BRICKS8
        nbytes = st["page_reads"] * VOXEL_PAGE_ENTRY_BYTES + st["cell_reads"] * VOXEL_CELL_BYTES
        launches = -(-cfg.samples // 8) if cfg.samples > 16 else 1
        achieved = nbytes / (ms / 1e3) / 1e9
        traffic = None
        if (cfg.width, cfg.height, cfg.samples, cfg.voxel_resolution) == VOXEL_TRAFFIC_WORKLOAD:
            traffic = 1024 * (VOXEL_TRAFFIC_FETCH_KIB * TRAFFIC_FETCH_SCALE +
                              VOXEL_TRAFFIC_WRITE_KIB * TRAFFIC_WRITE_SCALE)
        line["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                            "kernel": "voxel_persistent_kernel",
                            "launches_per_step": launches, "bytes_per_step": nbytes}
        line["config"]["counts"] = st
        if not args.no_cpu_baseline:
            # The FP32 oracle's voxel path tracer over the GPU's own tables
            # (read back), on a band of rows: the same traversal and shading;
            # pixels differ only where a path's direction, drawn with libm's
            # trigonometry on the CPU, rounds differently and takes another
            # cell.
            from oracle import bind
            threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
            img = accum.numpy().reshape(cfg.height, cfg.width, 4)
            tables = api.voxel_build(cfg)
            y0 = cfg.height // 2 - args.cpu_rows * 4
            y1 = y0 + args.cpu_rows * 8
            t = time.perf_counter()
            band = bind.voxel_render_tables(cfg, tables, threads=threads, rows=(y0, y1))[y0:y1]
            dt = time.perf_counter() - t
            n = (y1 - y0) * cfg.width * cfg.samples
            line["cpu_baseline"] = {"value": n / dt, "unit": "path samples/s", "cores": threads, "kind": "port",
                                    "sample": f"FP32 C++ oracle voxel path tracer (OpenMP) over the GPU-built "
                                              f"tables read back, rows {y0}-{y1 - 1} at {cfg.samples} spp "
                                              f"({n} path samples, {dt:.1f} s)",
                                    "cpu_model": cpu_model(),
                                    "band_equal_pixel_fraction": float(np.mean(np.all(
                                        band[..., :3].astype(np.float32) == img[y0:y1, :, :3], axis=-1))),
                                    "band_max_abs_diff": float(np.abs(band[..., :3] - img[y0:y1, :, :3]).max())}
    plan.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
