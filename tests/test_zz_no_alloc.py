"""The allocation-free surrogate gradient (lsdem_surrogate_gradient_noalloc) against the
allocating entry point lsdem_surrogate_gradient: identical gradients and loss, bit for bit, on two
calls in a row (the second with fewer samples, so it runs on what the first left in the shared
scratch), plus its capacity and shape-id errors.  Inputs and outputs are device buffers allocated
through the HIP runtime the library itself runs on (no second runtime in the process); the calls go
through the C ABI."""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import surrogate_ref as R  # noqa: E402
from hip_rt import Hip  # noqa: E402

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


def _noalloc(sim, hip, p, shape, t, n_param):
    g0 = np.full(n_param, np.nan, np.float32)
    l0 = np.full(1, np.nan, np.float64)
    dp, ds, dt, dg, dl = (hip.put(x) for x in (p, shape, t, g0, l0))
    try:
        rc = sim.lib.lsdem_surrogate_gradient_noalloc(sim.ctx, C.c_int64(len(shape)), dp, ds, dt, dg, dl)
        return rc, hip.get(dg, g0), float(hip.get(dl, l0)[0])
    finally:
        hip.free(dp, ds, dt, dg, dl)


def test_noalloc_gradient_equals_allocating_entry_point(L, S):
    sim, half = _setup(L, S)
    hip = Hip()
    n_param = L.SURROGATE_N_PARAM + 16 * 3
    rng = np.random.default_rng(5)
    for n in (4096, 1280):  # the second call is smaller than the first
        p, shape, t = _samples(rng, n, half)
        loss_ref, g_ref = sim.surrogate_gradient(p, shape, t, 3)
        rc, g, loss = _noalloc(sim, hip, p, shape, t, n_param)
        print(f"n={n}: rc={rc} loss {loss!r} vs {loss_ref!r}", flush=True)
        assert rc == 0, sim._err()
        assert np.array_equal(g.view(np.uint32), g_ref.view(np.uint32)), f"n={n}: gradient differs"
        assert loss == loss_ref


def test_noalloc_errors(L, S):
    sim, half = _setup(L, S)
    hip = Hip()
    n_param = L.SURROGATE_N_PARAM + 16 * 3
    p, shape, t = _samples(np.random.default_rng(9), 64 * 400, half)  # beyond the 16 MiB scratch
    rc, _, _ = _noalloc(sim, hip, p, shape, t, n_param)
    assert rc == L.LSDEM_ERR_CAPACITY and "scratch" in sim._err()
    p, shape, t = _samples(np.random.default_rng(9), 128, half)
    shape[17] = 7  # no such shape in the family
    rc, _, _ = _noalloc(sim, hip, p, shape, t, n_param)
    assert rc == L.LSDEM_ERR_ARG and "family" in sim._err()
    # a valid call afterwards still equals the allocating entry point
    shape[17] = 1
    loss_ref, g_ref = sim.surrogate_gradient(p, shape, t, 3)
    rc, g, loss = _noalloc(sim, hip, p, shape, t, n_param)
    assert rc == 0 and np.array_equal(g.view(np.uint32), g_ref.view(np.uint32)) and loss == loss_ref
