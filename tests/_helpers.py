"""Shared comparison helpers for the parity tests."""
import numpy as np


def rel_close(got, want, rtol, atol=0.0):
    """Per-element |got - want| <= rtol |want| + atol; returns (ok, worst relative excess)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape, (got.shape, want.shape)
    err = np.abs(got - want)
    lim = rtol * np.abs(want) + atol
    worst = float(np.max(err / np.maximum(np.abs(want), 1e-300))) if err.size else 0.0
    return bool(np.all(err <= lim)), worst


def assert_rel(got, want, rtol, atol=0.0, what=""):
    ok, worst = rel_close(got, want, rtol, atol)
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    i = int(np.argmax(np.abs(got - want) - rtol * np.abs(want)))
    assert ok, (f"{what}: worst |got-want| exceeds {rtol:g}*|want|+{atol:g}; at flat index {i} got "
                f"{got.reshape(-1)[i]!r} want {want.reshape(-1)[i]!r}")
