"""GPU checks of the weight surrogate's kernels (muca_mlp_*): forward and
backward on MFMA (BF16 operands, FP32 accumulation) against the CPU
restatement in oracle/mlp_oracle.py, which applies the same BF16 rounding in
float64 and is itself pinned by finite differences (tests/test_mlp_oracle.py).
"""
import numpy as np
import pytest

import muca_amd as M
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = np.stack([rng.uniform(0, 1, n), rng.uniform(0.25, 2, n), rng.uniform(-0.4, 0.4, n),
                  rng.uniform(-0.5, 0.5, n)], 1).astype(np.float32)
    t = (np.sin(3 * x[:, 0]) * x[:, 1] - 0.3 * x[:, 2] + 0.1 * x[:, 3]).astype(np.float32)
    return x, t


@pytest.mark.parametrize("n", [64, 200, 1024])
def test_forward_matches_oracle(n):
    p = MO.init_params(1)
    x, _ = _data(n, 2)
    with M.Mlp() as net:
        net.set_params(p)
        y = net.predict(x)
    yr, _, _ = MO.forward_backward(p, x)
    # same BF16 operands; FP32 accumulation on the GPU vs float64 here, and an
    # occasional BF16 rounding flip of an activation near a tie
    assert np.max(np.abs(y - yr)) <= 2e-3 * max(1.0, np.max(np.abs(yr)))
    assert np.sqrt(np.mean((y - yr) ** 2)) <= 2e-4 * max(1.0, np.max(np.abs(yr)))


@pytest.mark.parametrize("n", [64, 333, 2048])
def test_gradient_matches_oracle(n):
    p = MO.init_params(3)
    x, t = _data(n, 4)
    with M.Mlp() as net:
        net.set_params(p)
        g, loss = net.gradient(x, t)
    _, lr, gr = MO.forward_backward(p, x, t)
    assert abs(loss - lr) <= 1e-4 * abs(lr)
    blocks = {"W1": (MO.OW1, MO.OB1), "b1": (MO.OB1, MO.OW2), "W2": (MO.OW2, MO.OB2),
              "b2": (MO.OB2, MO.OW3), "W3": (MO.OW3, MO.OB3), "b3": (MO.OB3, MO.OW4),
              "w4": (MO.OW4, MO.OB4), "b4": (MO.OB4, MO.NPARAMS)}
    for name, (a, b) in blocks.items():
        ga, gb = g[a:b].astype(np.float64), gr[a:b]
        err = np.linalg.norm(ga - gb) / max(np.linalg.norm(gb), 1e-30)
        assert err <= 5e-3, (name, err)


@pytest.mark.parametrize("weighted", [False, True])
def test_training_reduces_loss_like_oracle(weighted):
    """A few hundred Adam steps on the GPU follow the CPU schedule: the loss
    falls by the same order and the parameters stay close."""
    x, t = _data(4096, 5)
    w = np.random.default_rng(7).uniform(0.2, 3.0, len(x)).astype(np.float32) if weighted else None
    p0 = MO.init_params(6)
    with M.Mlp() as net:
        net.set_params(p0)
        _, l0 = net.gradient(x, t)
        net.set_params(p0)
        net.set_data(x, t, w)
        net.train(200, 1024, 3e-3, 3e-4)
        p = net.get_params()
        _, l1 = net.gradient(x, t)
    pr = MO.adam_train(p0, x.astype(np.float64), t.astype(np.float64), 200, 1024, 3e-3, 3e-4,
                       w=None if w is None else w.astype(np.float64))
    _, lr1, _ = MO.forward_backward(pr, x, t)
    assert l1 < 0.1 * l0
    assert l1 < 2.0 * lr1 and lr1 < 2.0 * l1
    assert np.linalg.norm(p - pr) <= 0.05 * np.linalg.norm(pr)
