"""FP64 numpy restatement of the learned culling surrogate (a Fourier-feature MLP per shape family).

TEST INFRASTRUCTURE ONLY: the checker the GPU training and inference kernels are compared against;
nothing in the product imports it.  The surrogate has no counterpart in the reference (it is the
north-star addition); the network is:

    u      = p * inv_len                          (p: unscaled local point of the shape, 3-d)
    x[0:96]  : for axis a, frequency j (0..15): x[32a + 2j] = sin(w_j u_a), x[32a + 2j + 1] = cos(w_j u_a),
               w_j = pi * 2^(j/3)
    x[96:112]: the shape's 16-d embedding
    h1 = silu(x W1 + b1)   W1: 112 x 64
    h2 = silu(h1 W2 + b2), h3 = silu(h2 W3 + b3), h4 = silu(h3 W4 + b4)   W2..W4: 64 x 64
    out = h4 . w5 + b5,   predicted SDF (unscaled units) = out / inv_len

Rounding points of the GPU kernels (BF16 operands of the matrix products, FP32 accumulation) can be
mimicked with bf16=True so that gradients agree to accumulation order.
"""
from __future__ import annotations

import numpy as np

N_FREQ = 16
N_EMB = 16
N_IN = 6 * N_FREQ + N_EMB  # 112
N_HID = 64
FREQS = np.pi * 2.0 ** (np.arange(N_FREQ) / 3.0)


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def features(p: np.ndarray, emb: np.ndarray, inv_len: float) -> np.ndarray:
    """(n, 112) inputs: Fourier features of p * inv_len, then the per-sample embedding rows emb (n, 16)."""
    u = np.asarray(p, float) * inv_len
    x = np.empty((len(u), N_IN))
    for a in range(3):
        t = u[:, a:a + 1] * FREQS[None, :]
        x[:, 32 * a:32 * a + 32:2] = np.sin(t)
        x[:, 32 * a + 1:32 * a + 32:2] = np.cos(t)
    x[:, 96:] = emb
    return x


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(rng: np.random.Generator, n_shapes: int) -> dict:
    """He-style initialisation (the GPU side is given these values, not its own)."""
    P = {"W1": rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_HID)), "b1": np.zeros(N_HID)}
    for k in (2, 3, 4):
        P[f"W{k}"] = rng.normal(0, np.sqrt(2.0 / N_HID), (N_HID, N_HID))
        P[f"b{k}"] = np.zeros(N_HID)
    P["w5"] = rng.normal(0, np.sqrt(1.0 / N_HID), N_HID)
    P["b5"] = np.zeros(1)
    P["emb"] = rng.normal(0, 0.5, (n_shapes, N_EMB))
    return P


def forward(P: dict, p: np.ndarray, shape: np.ndarray, inv_len: float, bf=False):
    """Returns out (n,) and the cache for backward."""
    r = bf16 if bf else (lambda a: a)
    x = features(f32(p) if bf else p, P["emb"][shape], inv_len)
    if bf:
        x = f32(x)
    xs = r(x)
    zs, hs = [], [xs]
    h = xs
    for k in (1, 2, 3, 4):
        W = r(P[f"W{k}"])
        z = h @ W + P[f"b{k}"]
        a = silu(z)
        zs.append(z)
        h = r(a) if k < 4 else a
        hs.append(h)
    out = h @ P["w5"] + P["b5"][0]
    return out, {"x": xs, "zs": zs, "hs": hs, "shape": shape}


def loss_and_grads(P: dict, p, shape, target, inv_len: float, bf=False):
    """Mean squared error of out against target (normalised units) and its gradients."""
    r = bf16 if bf else (lambda a: a)
    out, c = forward(P, p, shape, inv_len, bf)
    n = len(out)
    err = out - target
    loss = float(np.mean(err * err))
    dout = 2.0 * err / n
    G = {"w5": c["hs"][4].T @ dout, "b5": np.array([dout.sum()])}
    dh = np.outer(dout, P["w5"])
    for k in (4, 3, 2, 1):
        dz = dh * dsilu(c["zs"][k - 1])
        G[f"b{k}"] = dz.sum(0)
        dzr = r(dz)
        G[f"W{k}"] = c["hs"][k - 1].T @ dzr
        dh = dzr @ r(P[f"W{k}"]).T
    G["emb"] = np.zeros_like(P["emb"])
    np.add.at(G["emb"], c["shape"], dh[:, 96:])
    return loss, G, out


def adam(P: dict, G: dict, state: dict, lr: float, b1=0.9, b2=0.999, eps=1e-8):
    t = state["t"] = state.get("t", 0) + 1
    for k in P:
        m = state.setdefault("m_" + k, np.zeros_like(P[k]))
        v = state.setdefault("v_" + k, np.zeros_like(P[k]))
        m[...] = b1 * m + (1 - b1) * G[k]
        v[...] = b2 * v + (1 - b2) * G[k] * G[k]
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        P[k] = P[k] - lr * mh / (np.sqrt(vh) + eps)
