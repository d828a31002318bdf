"""The tuner's bookkeeping on the CPU, with a stand-in simulation whose time per step is set by its
configuration: the fastest setting is chosen, each size bucket is tuned on its largest member, and
the skin is derived from the shapes' node radius and the largest scale."""
import time
from types import SimpleNamespace

import numpy as np


class _FakeSim:
    def __init__(self, cfg):
        self.cfg = cfg

    def step(self, n):
        # narrow_batch 1 is fastest; the others cost 5x more per step
        time.sleep(n * (2e-5 if self.cfg["narrow_batch"] == 1 else 1e-4))


class _FakeScene:
    def __init__(self, n, r=0.5):
        self.cfg = {"dt": 1e-5}
        self.shapes = [{"nodes": np.array([[r, 0.0, 0.0], [0.0, 0.0, -2 * r]])}, {"nodes": np.zeros((0, 3))}]
        self.particles = {"scale": np.array([1.0, 3.0])}
        self.n_particles = n
        self.loaded = []

    def load(self, sim):
        self.loaded.append(dict(sim.cfg))
        return sim


def test_tune_picks_fastest():
    import lsdem_amd.autotune as A
    L = SimpleNamespace(Sim=_FakeSim)
    sc = _FakeScene(100)
    space = {"narrow_batch": (0, 1, 4), "skin_factor": (0.1,)}
    best, times = A.tune(L, sc, {"broad_phase_launches": 1}, space, warmup=1, steps=40)
    assert best == {"narrow_batch": 1, "skin_factor": 0.1}
    assert len(times) == 3
    for cfg in sc.loaded:
        assert cfg["broad_phase_launches"] == 1 and cfg["dt"] == 1e-5
        assert np.isclose(cfg["skin"], 0.1 * 1.0 * 3.0)  # node radius 1.0 (2r) times the largest scale


def test_buckets_tuned_on_largest_member():
    import lsdem_amd.autotune as A
    L = SimpleNamespace(Sim=_FakeSim)
    scs = [_FakeScene(n) for n in (100, 250, 400, 900, 350)]
    chosen, report = A.tune_buckets(L, scs, edges=(300, 600), space={"narrow_batch": (0, 1)}, warmup=1, steps=20)
    assert sorted(report) == [0, 1, 2]
    assert report[0]["representative"] == 1 and report[1]["representative"] == 2 and report[2]["representative"] == 3
    assert all(c == {"narrow_batch": 1} for c in chosen)
    assert len(scs[1].loaded) == 2 and scs[0].loaded == [] and scs[4].loaded == []
    cfg = A.apply(scs[0], {"block_size": 64}, chosen[0])
    assert cfg["narrow_batch"] == 1 and cfg["block_size"] == 64
