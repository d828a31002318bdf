"""The allocation-free surrogate gradient (lsdem_surrogate_gradient_noalloc) against the
allocating entry point lsdem_surrogate_gradient: identical gradients and loss, bit for bit, on two
calls in a row (the second with fewer samples, so it runs on what the first left in the shared
scratch), plus its capacity and shape-id errors.  Inputs and outputs are device buffers (torch
tensors on the context's device); the calls go through the C ABI."""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import surrogate_ref as R  # noqa: E402

pytestmark = pytest.mark.gpu


def _flatten(P):
    return np.concatenate([P["W1"].ravel(), P["b1"], P["W2"].ravel(), P["b2"], P["W3"].ravel(), P["b3"],
                           P["W4"].ravel(), P["b4"], P["w5"], P["b5"], P["emb"].ravel()]).astype(np.float32)


def _setup(L, S):
    sc = S.scene_c2(n=600, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.06, gap=-0.15)
    sim = sc.load(L.Sim(sc.cfg))
    P = R.init_params(np.random.default_rng(11), 3)
    sim.surrogate_set_weights(_flatten(P))
    half = max(s["spacing"] * (s["sdf"].shape[0] - 1) * 0.5 for s in sc.shapes[:3])
    return sim, half


def _samples(rng, n, half):
    p = (rng.uniform(-1.0, 1.0, (n, 3)) * half).astype(np.float32)
    shape = rng.integers(0, 3, n).astype(np.int32)
    t = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    return p, shape, t


def _noalloc(sim, torch, p, shape, t, n_param):
    dev = torch.device("cuda", 0)
    dp = torch.from_numpy(p).to(dev)
    ds = torch.from_numpy(shape).to(dev)
    dt = torch.from_numpy(t).to(dev)
    dg = torch.full((n_param,), float("nan"), dtype=torch.float32, device=dev)
    dl = torch.full((1,), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    rc = sim.lib.lsdem_surrogate_gradient_noalloc(sim.ctx, C.c_int64(len(shape)), C.c_void_p(dp.data_ptr()),
                                                  C.c_void_p(ds.data_ptr()), C.c_void_p(dt.data_ptr()),
                                                  C.c_void_p(dg.data_ptr()), C.c_void_p(dl.data_ptr()))
    return rc, dg.cpu().numpy(), float(dl.cpu().numpy()[0])


def test_noalloc_gradient_equals_allocating_entry_point(L, S):
    import torch
    sim, half = _setup(L, S)
    n_param = L.SURROGATE_N_PARAM + 16 * 3
    rng = np.random.default_rng(5)
    for n in (4096, 1280):  # the second call is smaller than the first
        p, shape, t = _samples(rng, n, half)
        loss_ref, g_ref = sim.surrogate_gradient(p, shape, t, 3)
        rc, g, loss = _noalloc(sim, torch, p, shape, t, n_param)
        print(f"n={n}: rc={rc} loss {loss!r} vs {loss_ref!r}", flush=True)
        assert rc == 0, sim._err()
        assert np.array_equal(g.view(np.uint32), g_ref.view(np.uint32)), f"n={n}: gradient differs"
        assert loss == loss_ref


def test_noalloc_errors(L, S):
    import torch
    sim, half = _setup(L, S)
    n_param = L.SURROGATE_N_PARAM + 16 * 3
    p, shape, t = _samples(np.random.default_rng(9), 64 * 400, half)  # beyond the 16 MiB scratch
    rc, _, _ = _noalloc(sim, torch, p, shape, t, n_param)
    assert rc == L.LSDEM_ERR_CAPACITY and "scratch" in sim._err()
    p, shape, t = _samples(np.random.default_rng(9), 128, half)
    shape[17] = 7  # no such shape in the family
    rc, _, _ = _noalloc(sim, torch, p, shape, t, n_param)
    assert rc == L.LSDEM_ERR_ARG and "family" in sim._err()
    # a valid call afterwards still equals the allocating entry point
    shape[17] = 1
    loss_ref, g_ref = sim.surrogate_gradient(p, shape, t, 3)
    rc, g, loss = _noalloc(sim, torch, p, shape, t, n_param)
    assert rc == 0 and np.array_equal(g.view(np.uint32), g_ref.view(np.uint32)) and loss == loss_ref
