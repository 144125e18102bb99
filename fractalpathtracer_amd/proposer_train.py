"""Proposer training (PyTorch), run as its own process by proposer.train().

Input: (n, 4) float32 samples (x, y, z, exact distance) and an initial
parameter vector; output: the fitted FPT_PROPOSER_PARAM_COUNT floats in a
.npz weight file.  The network is the one the GPU evaluates (features and
layer layout in include/fpt_abi.h); the last 20% of the steps train through
the inference rounding (BF16 weights and hidden-layer inputs, with
straight-through gradients) so the FP32 fit and the BF16 network agree.

Loss: |q - d| + 2 relu(q - d) on the distance clamped to [-0.25, 1.0], i.e.
over-estimates (which make a far step overshoot) cost three times as much as
under-estimates.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np
import torch

from fractalpathtracer_amd import proposer as P


def features(v: torch.Tensor) -> torch.Tensor:
    """(n, 3) normalised points -> (n, 128) features, index 32a + 8q + 2i + c."""
    n = v.shape[0]
    f = torch.zeros((n, P.INPUTS), dtype=v.dtype, device=v.device)
    for a in range(3):
        for q in range(4):
            for i in range(4):
                w = math.pi * 2.0 ** ((4 * q + i) / 2.0)
                k = 32 * a + 8 * q + 2 * i
                f[:, k] = torch.sin(w * v[:, a])
                f[:, k + 1] = torch.cos(w * v[:, a])
    f[:, 96:99] = v
    return f


def _bf16(x: torch.Tensor) -> torch.Tensor:
    """Round to BF16 (nearest even) with a straight-through gradient."""
    return x + (x.to(torch.bfloat16).to(x.dtype) - x).detach()


def forward(params: dict, v: torch.Tensor, quantised: bool) -> torch.Tensor:
    q = _bf16 if quantised else (lambda t: t)
    h = features(v)
    for i in (1, 2, 3):
        h = torch.nn.functional.silu(q(h) @ q(params[f"W{i}"]).T + params[f"b{i}"])
    h = torch.nn.functional.silu(q(h) @ q(params["W4"]).T + params["b4"])
    return h @ params["w5"] + params["b5"][0]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", required=True)
    ap.add_argument("--init", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--box-half", type=float, default=P.BOX_HALF)
    ap.add_argument("--device", default="auto")
    a = ap.parse_args()

    dev = a.device
    if dev == "auto":
        dev = "cuda" if torch.cuda.is_available() else "cpu"
    torch.manual_seed(a.seed)
    data = torch.from_numpy(np.load(a.samples)).to(dev)
    v = data[:, :3] / a.box_half
    d = data[:, 3].clamp(-0.25, 1.0)
    init = P.split(np.load(a.init))
    params = {k: torch.tensor(x, device=dev, requires_grad=True) for k, x in init.items()}
    opt = torch.optim.Adam(params.values(), lr=a.lr)
    n = data.shape[0]
    qat_from = int(a.steps * 0.8)
    gen = torch.Generator(device=dev)
    gen.manual_seed(a.seed)
    t0 = time.perf_counter()
    for step in range(a.steps):
        lr = a.lr * 0.5 * (1.0 + math.cos(math.pi * step / a.steps))
        for g in opt.param_groups:
            g["lr"] = lr
        idx = torch.randint(0, n, (min(a.batch, n),), device=dev, generator=gen)
        pred = forward(params, v[idx], quantised=step >= qat_from)
        err = pred - d[idx]
        loss = (err.abs() + 2.0 * torch.relu(err)).mean()
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        if step % 500 == 0 or step == a.steps - 1:
            print(f"step {step} loss {loss.item():.5f} lr {lr:.2e}", flush=True)
    with torch.no_grad():
        pred = torch.cat([forward(params, v[i:i + 262144], True) for i in range(0, n, 262144)])
        err = (pred - data[:, 3]).cpu().numpy()
        dd = data[:, 3].cpu().numpy()
    far = dd > 0.05
    meta = {
        "steps": a.steps, "batch": a.batch, "lr": a.lr, "device": dev,
        "train_loop_seconds": time.perf_counter() - t0,
        "mae_all": float(np.mean(np.abs(err))),
        "mae_far": float(np.mean(np.abs(err[far]))) if far.any() else 0.0,
        # far step 0.5 q overshoots the surface iff q > 2 d, i.e. err > d
        "overshoot_far_frac": float(np.mean(err[far] > dd[far])) if far.any() else 0.0,
    }
    out = P.join({k: t.detach().cpu().numpy() for k, t in params.items()})
    P.save(a.out, out, meta)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
