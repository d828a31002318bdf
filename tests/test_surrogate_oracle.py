"""Pins the FP64 surrogate restatement (oracle/surrogate_ref.py), which the GPU kernels are checked
against: gradients against central finite differences of its own loss, the feature map against
closed forms, the BF16 rounding helper against exact cases, Adam against its first-step closed form."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import surrogate_ref as R  # noqa: E402


def _problem(seed=0, n=37, shapes=3):
    rng = np.random.default_rng(seed)
    P = R.init_params(rng, shapes)
    p = rng.uniform(-0.8, 0.8, (n, 3))
    s = rng.integers(0, shapes, n)
    t = rng.normal(0, 0.3, n)
    return P, p, s, t


def test_features_closed_form():
    p = np.array([[0.0, 0.5, -0.25]])
    emb = np.arange(16.0)[None, :]
    x = R.features(p, emb, 2.0)
    u = np.array([0.0, 1.0, -0.5])
    for a in range(3):
        for j in range(16):
            w = np.pi * 2.0 ** (j / 3.0)
            assert x[0, 32 * a + 2 * j] == pytest.approx(np.sin(w * u[a]), abs=1e-15)
            assert x[0, 32 * a + 2 * j + 1] == pytest.approx(np.cos(w * u[a]), abs=1e-15)
    assert np.array_equal(x[0, 96:], emb[0])
    assert R.FREQS[3] == pytest.approx(2 * np.pi)  # 2^(3/3) pi


def test_bf16_rounding():
    assert R.bf16(np.array([1.0]))[0] == 1.0
    # 1 + 2^-8 lies halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    assert R.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    assert R.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert R.bf16(np.array([-3.0]))[0] == -3.0


@pytest.mark.parametrize("name", ["W1", "b1", "W2", "b3", "W4", "b4", "w5", "b5", "emb"])
def test_gradients_match_finite_differences(name):
    P, p, s, t = _problem()
    _, G, _ = R.loss_and_grads(P, p, s, t, 1.3)
    rng = np.random.default_rng(1)
    flat = P[name].reshape(-1)
    idx = rng.choice(flat.size, size=min(6, flat.size), replace=False)
    if name == "emb":
        idx = np.array([s[0] * 16 + 3, s[1] * 16 + 7])  # rows that occur in the batch
    for i in idx:
        h = 1e-6
        old = flat[i]
        flat[i] = old + h
        lp, _, _ = R.loss_and_grads(P, p, s, t, 1.3)
        flat[i] = old - h
        lm, _, _ = R.loss_and_grads(P, p, s, t, 1.3)
        flat[i] = old
        fd = (lp - lm) / (2 * h)
        assert G[name].reshape(-1)[i] == pytest.approx(fd, rel=2e-5, abs=1e-9)


def test_bf16_gradients_close_to_fp64():
    P, p, s, t = _problem(n=200)
    _, G, _ = R.loss_and_grads(P, p, s, t, 1.3)
    _, Gb, _ = R.loss_and_grads(P, p, s, t, 1.3, bf=True)
    for k in G:
        e = np.linalg.norm(G[k] - Gb[k]) / (np.linalg.norm(G[k]) + 1e-30)
        assert e < 3e-2, (k, e)


def test_adam_first_step():
    P = {"a": np.array([1.0, -2.0])}
    G = {"a": np.array([0.5, -4.0])}
    st = {}
    R.adam(P, G, st, lr=0.1)
    # first step: m/(1-b1) = g, v/(1-b2) = g^2 -> update = lr * g / (|g| + eps)
    assert np.allclose(P["a"], [1.0 - 0.1, -2.0 + 0.1], atol=1e-7)
