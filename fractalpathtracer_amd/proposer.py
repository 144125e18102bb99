"""Learned far-field step proposer: parameter layout, weight files, training.

The network and how the march uses it are specified in include/fpt_abi.h
(FPT_PROPOSER_PARAM_COUNT, FptProposerSettings) and implemented on the GPU in
csrc/learned.h.  This module holds the host side:

  layout      offsets of W1 b1 W2 b2 W3 b3 W4 b4 w5 b5 in the flat float32
              parameter vector the C ABI takes;
  save/load   .npz weight files (no pickles);
  train       GPU training samples from the library's sampler, then the
              optimisation in a separate Python process (proposer_train.py,
              PyTorch), so the PyTorch HIP runtime never shares a process
              with libfpt_hip.so.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

INPUTS = 128
HIDDEN = 64
PARAM_COUNT = HIDDEN * INPUTS + HIDDEN + 3 * (HIDDEN * HIDDEN + HIDDEN) + HIDDEN + 1  # 20801

# Defaults of the learned-step workload (march settings in FptProposerSettings).
BOX_HALF = 1.5
BAND = 0.05
SWITCH_DISTANCE = 0.02
RESUME_DISTANCE = 0.1
STEP_SCALE = 0.5

ASSETS = Path(__file__).resolve().parent / "assets"


def layout() -> dict[str, tuple[int, tuple[int, ...]]]:
    """name -> (offset, shape) in the flat parameter vector."""
    out, off = {}, 0
    shapes = [("W1", (HIDDEN, INPUTS)), ("b1", (HIDDEN,))]
    for i in (2, 3, 4):
        shapes += [(f"W{i}", (HIDDEN, HIDDEN)), (f"b{i}", (HIDDEN,))]
    shapes += [("w5", (HIDDEN,)), ("b5", (1,))]
    for name, shape in shapes:
        out[name] = (off, shape)
        off += int(np.prod(shape))
    assert off == PARAM_COUNT
    return out


def split(params: np.ndarray) -> dict[str, np.ndarray]:
    p = np.asarray(params, np.float32).ravel()
    if p.size != PARAM_COUNT:
        raise ValueError(f"expected {PARAM_COUNT} parameters, got {p.size}")
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in layout().items()}


def join(parts: dict[str, np.ndarray]) -> np.ndarray:
    p = np.zeros(PARAM_COUNT, np.float32)
    for k, (o, s) in layout().items():
        p[o:o + int(np.prod(s))] = np.asarray(parts[k], np.float32).reshape(-1)
    return p


def init_params(seed: int = 0, gain: float = 1.0) -> np.ndarray:
    """He-style random initialisation (used by tests and as the training
    start point)."""
    rng = np.random.default_rng(seed)
    parts = {}
    for k, (_, s) in layout().items():
        if k.startswith("W"):
            parts[k] = rng.normal(0.0, gain * np.sqrt(2.0 / s[1]), s)
        elif k == "w5":
            parts[k] = rng.normal(0.0, gain * np.sqrt(1.0 / HIDDEN), s)
        else:
            parts[k] = rng.normal(0.0, 0.05 * gain, s)
    return join(parts)


def save(path: os.PathLike, params: np.ndarray, meta: dict) -> None:
    np.savez(path, params=np.asarray(params, np.float32), meta=np.array(json.dumps(meta)))


def load(path: os.PathLike) -> tuple[np.ndarray, dict]:
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        meta = json.loads(str(z["meta"]))
    if params.size != PARAM_COUNT:
        raise ValueError(f"{path}: expected {PARAM_COUNT} parameters, got {params.size}")
    return params, meta


def asset_path(name: str = "mandelbulb_p8") -> Path:
    return ASSETS / f"proposer_{name}.npz"


def train(cfg, samples: int = 1 << 22, seed: int = 0, steps: int = 4000, batch: int = 65536,
          lr: float = 2e-3, device: str = "auto", box_half: float = BOX_HALF, band: float = BAND,
          log=None) -> tuple[np.ndarray, dict]:
    """Generate `samples` training points of `cfg`'s scene on the GPU
    (half uniform in the box, half within `band` of the surface) and fit the
    proposer in a child process.  Returns (params, info)."""
    from . import api
    t0 = time.perf_counter()
    pts, fails = api.proposer_samples(cfg, samples, seed=seed, box_half=box_half, band=band)
    t_sample = time.perf_counter() - t0
    with tempfile.TemporaryDirectory() as tmp:
        sp = Path(tmp) / "samples.npy"
        wp = Path(tmp) / "weights.npz"
        np.save(sp, pts)
        init = init_params(seed)
        ip = Path(tmp) / "init.npy"
        np.save(ip, init)
        cmd = [sys.executable, "-m", "fractalpathtracer_amd.proposer_train", "--samples", str(sp), "--init", str(ip),
               "--out", str(wp), "--steps", str(steps), "--batch", str(batch), "--lr", str(lr), "--seed", str(seed),
               "--box-half", str(box_half), "--device", device]
        env = dict(os.environ)
        root = str(Path(__file__).resolve().parent.parent)
        env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
        t1 = time.perf_counter()
        r = subprocess.run(cmd, env=env, capture_output=True, text=True)
        if log:
            log(r.stdout)
        if r.returncode != 0:
            raise RuntimeError(f"proposer training failed ({r.returncode}):\n{r.stderr[-4000:]}")
        params, meta = load(wp)
        t_train = time.perf_counter() - t1
    meta.update(samples=int(samples), sample_failures=fails, sample_seconds=t_sample, train_seconds=t_train,
                box_half=box_half, band=band, seed=seed)
    return params, meta


def main(argv=None) -> int:
    """python -m fractalpathtracer_amd.proposer train --out FILE [...]:
    train the Mandelbulb (power 8, 12 iterations) proposer."""
    import argparse
    ap = argparse.ArgumentParser(prog="python -m fractalpathtracer_amd.proposer")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--samples", type=int, default=1 << 22)
    t.add_argument("--steps", type=int, default=4000)
    t.add_argument("--batch", type=int, default=65536)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--device", default="auto")
    a = ap.parse_args(argv)
    from . import scenes
    cfg = scenes.learned_step_scene(64, 64, 1)
    params, meta = train(cfg, samples=a.samples, seed=a.seed, steps=a.steps, batch=a.batch, device=a.device,
                         log=lambda s: print(s, end="", flush=True))
    meta["scene"] = "Mandelbulb power 8, 12 iterations, bailout 2"
    save(a.out, params, meta)
    print(json.dumps(meta), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
