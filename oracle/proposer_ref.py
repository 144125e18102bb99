"""Reference evaluation of the learned step proposer (numpy).

ORACLE — test infrastructure only: the checker the GPU proposer
(fpt_hip_proposer_eval) is compared with, never a product path.  The network
has no counterpart in the reference; its definition is the one in
include/fpt_abi.h (FPT_PROPOSER_PARAM_COUNT).  Pins (tests/test_proposer.py):
BF16 rounding against hand-derived bit patterns (ties to even), features
against closed forms at v = 0 and v = 1, and the whole forward pass against
the PyTorch training definition (fractalpathtracer_amd/proposer_train.py) run
in a separate process.
"""
from __future__ import annotations

import numpy as np

INPUTS = 128
HIDDEN = 64


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def features(points: np.ndarray, box_half: float) -> np.ndarray:
    """(n, 3) points -> (n, 128) float64 features, index 32a + 8q + 2i + c."""
    v = np.asarray(points, np.float64) / box_half
    f = np.zeros((v.shape[0], INPUTS))
    for a in range(3):
        for q in range(4):
            for i in range(4):
                w = np.pi * 2.0 ** ((4 * q + i) / 2.0)
                k = 32 * a + 8 * q + 2 * i
                f[:, k] = np.sin(w * v[:, a])
                f[:, k + 1] = np.cos(w * v[:, a])
    f[:, 96:99] = v
    return f


def _silu(x):
    return x / (1.0 + np.exp(-x))


def forward(params: np.ndarray, points: np.ndarray, box_half: float) -> np.ndarray:
    """The BF16 inference network: weights and the inputs of W1..W4 rounded
    to BF16, sums and everything else in float64."""
    from fractalpathtracer_amd import proposer as P
    p = P.split(params)
    h = features(points, box_half)
    for i in (1, 2, 3, 4):
        W = bf16_round(p[f"W{i}"]).astype(np.float64)
        x = bf16_round(h.astype(np.float32)).astype(np.float64)
        h = _silu(x @ W.T + p[f"b{i}"].astype(np.float64))
    return h @ p["w5"].astype(np.float64) + float(p["b5"][0])
