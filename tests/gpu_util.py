"""Helpers shared by the GPU parity tests."""

import sys

import numpy as np
import torch


def to_np(t):
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy()
    return np.asarray(t)


def dev(a, dtype=torch.float32):
    return torch.as_tensor(np.asarray(a)).to("cuda", dtype)


def rel_err(got, ref):
    """max |got - ref| / max |ref| (relative L-infinity error)."""
    got, ref = to_np(got).astype(np.float64), np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    scale = max(np.abs(ref).max(), 1e-30)
    return float(np.abs(got - ref).max() / scale)


def assert_close(got, ref, tol, what=""):
    e = rel_err(got, ref)
    assert e <= tol, f"{what}: relative L-inf error {e:.3e} > {tol:.1e}"
    return e


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()
