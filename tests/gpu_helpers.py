"""Shared helpers of the GPU parity tests: numpy <-> torch transfer and
running the same case through the HIP library and the CPU oracle."""
from __future__ import annotations

import numpy as np
import torch

NP_OF = {torch.float64: np.float64, torch.float32: np.float32}


def to_dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda:0")


def to_host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def interior(a, ih=1):
    return a[ih:a.shape[0] - ih, ih:a.shape[1] - ih, ih:a.shape[2] - ih]


def bits_equal(a, b):
    """Bitwise equality of two float arrays, treating +0 and -0 as equal."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def describe_diff(a, b):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    idx = np.unravel_index(int(np.argmax(d)), d.shape)
    return f"max|diff|={d.max():.3e} at {idx} ({int((d > 0).sum())} cells differ)"
