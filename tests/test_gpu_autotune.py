"""Tuned settings change the time per step, never the results: small turning drums stepped under
every setting of the tuner's space end bit-identical to the default configuration, and a bucketed
tuning over a few drums returns settings from that space."""
import numpy as np
import pytest

from test_gpu_drum import _small_drum

pytestmark = pytest.mark.gpu

SPACE = {"block_size": (64, 256), "narrow_batch": (0, 1, 4), "skin_factor": (0.04, 0.16)}


def _state(sim):
    return sim.get_particles(), sim.get_walls(), sim.counters()["contacts"], [sim.contacts(l) for l in (0, 1)]


def test_every_setting_same_results(L, S):
    import lsdem_amd.autotune as A
    sc = _small_drum(S, seed=5, rpm=45.0)
    base = {"broad_phase_launches": 1}
    chunks = (1, 19, 40)
    ref = sc.load(L.Sim({**sc.cfg, **base}))
    want = []
    for n in chunks:
        ref.step(n)
        want.append(_state(ref))
    assert len(want[0][3][1]["master"]) > 0  # wall contacts after the first step
    for setting in A._settings(SPACE):
        sim = sc.load(L.Sim(A.apply(sc, base, setting)))
        for n, (pr, wr, nc, cr) in zip(chunks, want):
            sim.step(n)
            p, w, c, cl = _state(sim)
            for k in ("pos", "quat", "vel", "angvel", "force", "torque"):
                assert np.array_equal(p[k], pr[k]), f"{setting}: {k} differs"
            for k in w:
                assert np.array_equal(w[k], wr[k]), f"{setting}: wall {k} differs"
            assert c == nc
            for l in (0, 1):
                for k in cr[l]:
                    assert np.array_equal(cl[l][k], cr[l][k]), f"{setting}: contact list {l} {k} differs"


def test_tune_buckets_on_drums(L, S):
    import lsdem_amd.autotune as A
    scs = [_small_drum(S, seed=20 + k, rpm=10.0 + 10.0 * k) for k in range(3)]
    space = {"block_size": (64, 256), "narrow_batch": (0, 4)}
    chosen, report = A.tune_buckets(L, scs, edges=(1000,), base={"broad_phase_launches": 1}, space=space,
                                    warmup=5, steps=40)
    assert list(report) == [0]
    assert chosen[0] == chosen[1] == chosen[2] == report[0]["best"]
    assert chosen[0]["block_size"] in space["block_size"] and chosen[0]["narrow_batch"] in space["narrow_batch"]
    assert len(report[0]["ms_per_step"]) == 4 and all(v > 0 for v in report[0]["ms_per_step"].values())
