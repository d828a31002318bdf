"""The broad phase as one cooperative launch (grid barriers between its phases) against the same
phases as separate gated launches: identical candidate pairs, contacts and states, step by step,
through rebuilds and periodic wraps."""
import numpy as np
import pytest

from parity import LISTS

pytestmark = pytest.mark.gpu


def test_single_launch_matches_separate_launches(L, S):
    sc = S.scene_c2(n=1500, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.05, gap=-0.1, seed=17)
    rng = np.random.default_rng(5)
    sc.particles["vel"][:, :2] = rng.normal(0.0, 5.0, (sc.n_particles, 2))  # frequent rebuilds and wraps
    sc.cfg["skin"] = 1e-3
    one = sc.load(L.Sim(sc.cfg))
    sep = sc.load(L.Sim({**sc.cfg, "broad_phase_launches": 1}))
    one.profile_begin(60)
    for s in range(60):
        one.step(1)
        sep.step(1)
        assert one.counters()["pairs"] == sep.counters()["pairs"], f"step {s}: candidate pairs differ"
        for l in range(5):
            ca, cb = one.contacts(l), sep.contacts(l)
            for k in ca:
                assert np.array_equal(ca[k], cb[k]), f"step {s} list {LISTS[l]}: {k} differs"
        a, b = one.get_particles(), sep.get_particles()
        for k in ("pos", "quat", "vel", "angvel", "force", "torque"):
            assert np.array_equal(a[k], b[k]), f"step {s}: {k} differs"
    prof = one.profile_end()
    print("rebuilds", prof["rebuilds"], "pairs", prof["pairs"], "evaluated", prof["evaluated_pairs"])
    assert prof["rebuilds"] >= 5
