"""Pins the CPU oracle of the tile-cost network (oracle/mlp_ref.py):

* the float64 backward pass against central finite differences of the
  float64 loss, parameter by parameter on a sample of every block (a
  transposed weight, a dropped ReLU mask or a wrong sign fails it);
* the bf16 backward pass stays within bf16 rounding of the float64 one;
* bf16 rounding against exhaustive expectations on representative codes.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import mlp_ref as M  # noqa: E402


def data(n=96, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 2, size=(n, M.F)).astype(np.float32)
    t = rng.uniform(5, 20, size=n).astype(np.float32)
    return x, t


def loss64(p, x, t):
    y, _ = M.forward(p, x, bf16_mode=False)
    return float(((y - t.astype(np.float64)) ** 2).mean())


@pytest.mark.parametrize("block", ["W1", "b1", "W2", "b2", "w3", "b3"])
def test_float64_gradient_matches_finite_differences(block):
    x, t = data()
    p = M.init(1).astype(np.float64)
    g, _ = M.gradients(p, x, t, bf16_mode=False)
    lo, hi = {"W1": (M.O_W1, M.O_B1), "b1": (M.O_B1, M.O_W2), "W2": (M.O_W2, M.O_B2), "b2": (M.O_B2, M.O_W3),
              "w3": (M.O_W3, M.O_B3), "b3": (M.O_B3, M.O_B3 + 1)}[block]
    rng = np.random.default_rng(2)
    idx = rng.choice(np.arange(lo, hi), size=min(24, hi - lo), replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (loss64(pp, x, t) - loss64(pm, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (block, i, fd, g[i])


def test_bf16_gradient_close_to_float64():
    x, t = data(256, 3)
    p = M.init(4)
    g16, l16 = M.gradients(p, x, t, bf16_mode=True)
    g64, l64 = M.gradients(p.astype(np.float64), x, t, bf16_mode=False)
    assert abs(l16 - l64) <= 0.02 * l64
    # bf16 has 8 significant bits: agreement to a few percent of the block norm
    for lo, hi in [(M.O_W1, M.O_B1), (M.O_W2, M.O_B2), (M.O_W3, M.O_B3), (M.O_B1, M.O_W2), (M.O_B2, M.O_W3)]:
        err = np.linalg.norm(g16[lo:hi] - g64[lo:hi])
        assert err <= 0.05 * np.linalg.norm(g64[lo:hi]) + 1e-6, (lo, err)


def test_bf16_rounding():
    e = 2.0 ** 100
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -8, -2.5, e * (1 + 2 ** -8),
                     e * (1 + 2 ** -8 + 2 ** -20), 0.0], np.float32)
    # bf16 keeps 8 significant bits: ties go to the even mantissa
    want = np.array([1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -6, -2.5, e, e * (1 + 2 ** -7), 0.0], np.float32)
    assert np.array_equal(M.bf16(vals), want)
    assert np.isnan(M.bf16(np.array([np.nan], np.float32)))[0]


def test_forward_modes_agree_within_bf16():
    x, t = data(64, 5)
    p = M.init(6)
    y16, _ = M.forward(p, x, True)
    y64, _ = M.forward(p.astype(np.float64), x, False)
    assert np.allclose(y16, y64, rtol=0.03, atol=0.05)
