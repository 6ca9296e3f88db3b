"""The tile-cost network's MFMA kernels against the CPU oracle of the same
network (oracle/mlp_ref.py, bf16 operands, fp32 sums): forward outputs,
gradients of every parameter block, and Adam training on the device."""
import sys
from pathlib import Path

import numpy as np
import pytest

import cpvulkan_amd

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import mlp_ref as R  # noqa: E402

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 3, size=(n, 12)).astype(np.float32)
    t = (5 + 2 * np.log2(1 + np.abs(x[:, 0] * x[:, 2])) + x[:, 1]).astype(np.float32)
    return x, t


@pytest.fixture
def mlp(gpu_ctx):
    m = cpvulkan_amd.Mlp(gpu_ctx, seed=0)
    yield m
    m.close()


def test_params_roundtrip(mlp):
    p = R.init(7)
    mlp.set_params(p)
    assert np.array_equal(mlp.params(), p)


@pytest.mark.parametrize("rows", [1, 16, 37, 1000])
def test_forward_matches_oracle(mlp, rows):
    p = R.init(11)
    mlp.set_params(p)
    x, _ = _data(rows, rows)
    y = mlp.predict(x)
    want, _ = R.forward(p, x, bf16_mode=True)
    print(f"[mlp] forward rows={rows} max|d|={np.max(np.abs(y - want)):.3g}", flush=True)
    assert np.allclose(y, want, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("rows", [16, 333, 4096])
def test_gradients_match_oracle(mlp, rows):
    p = R.init(12)
    mlp.set_params(p)
    x, t = _data(rows, 100 + rows)
    g, loss = mlp.gradients(x, t)
    gw, lw = R.gradients(p, x, t, bf16_mode=True)
    assert abs(loss - lw) <= 1e-4 * max(1.0, lw)
    for name, lo, hi in [("W1", R.O_W1, R.O_B1), ("b1", R.O_B1, R.O_W2), ("W2", R.O_W2, R.O_B2), ("b2", R.O_B2, R.O_W3),
                         ("w3", R.O_W3, R.O_B3), ("b3", R.O_B3, R.O_B3 + 1)]:
        err = np.max(np.abs(g[lo:hi] - gw[lo:hi]))
        scale = max(np.max(np.abs(gw[lo:hi])), 1e-6)
        print(f"[mlp] grad {name} rows={rows}: max err {err:.3g} (scale {scale:.3g})", flush=True)
        assert err <= 2e-3 * scale, name


def test_training_reduces_loss(mlp):
    x, t = _data(20000, 5)
    _, before = mlp.gradients(x, t)
    after = mlp.train(x, t, steps=400, batch=1024, lr=3e-3, seed=1)
    print(f"[mlp] train: loss {before:.4f} -> {after:.4f}", flush=True)
    assert after < 0.1 * before
    # the trained network still matches the oracle's forward pass
    p = mlp.params()
    y = mlp.predict(x[:500])
    want, _ = R.forward(p, x[:500], bf16_mode=True)
    assert np.allclose(y, want, rtol=1e-4, atol=1e-4)
