"""Drum ensembles in one simulation on the GPU: every drum's contacts and forces against the oracle
step by step, every drum bit-identical to that drum simulated alone on the GPU, and graph replays of
the whole ensemble bit-identical to eager steps."""
import numpy as np
import pytest

from ensemble_scenes import COUNTS, small_ensemble
from oracle_binding import oracle_sim
from parity import step_and_compare

pytestmark = pytest.mark.gpu

KEYS = ("pos", "quat", "vel", "angvel", "force", "torque")


def test_ensemble_matches_oracle(L, S):
    sc = small_ensemble(S)
    gpu = sc.load(L.Sim(sc.cfg))
    ref = sc.load(oracle_sim(sc.cfg))
    counts, _ = step_and_compare(gpu, ref, 40, lists=(0, 1), ctol=1e-9)
    assert sum(c["pw"] for c in counts) > 0 and sum(c["pp"] for c in counts) > 0
    assert np.array_equal(gpu.get_walls()["quat"], ref.get_walls()["quat"])  # every drum turned the same way


def test_ensemble_at_configuration_parameters(L, S):
    """Six drums at the configuration's own sizes (radius 0.1 m, depth 0.05 m, SDF 32^3, 300 nodes),
    counts lognormal (median 4,000, within 500-20,000) scaled so 20,000 fills half the drum's lattice,
    speeds, friction and restitution drawn per drum; 20 steps against the oracle."""
    rng = np.random.default_rng(17)
    K = 6
    counts = np.clip(rng.lognormal(np.log(4000.0), 0.8, K), 500, 20000)
    sc = S.scene_drum_ensemble(counts, rng.uniform(1, 30, K), rng.uniform(0.1, 0.8, K), rng.uniform(0.3, 0.9, K),
                               seed=23, max_count=20000)
    assert len(set(np.bincount(sc.meta["drum_of"]))) == K  # uneven sizes
    gpu = sc.load(L.Sim(sc.cfg))
    ref = sc.load(oracle_sim(sc.cfg, threads=8))
    step_and_compare(gpu, ref, 20, lists=(0, 1), ctol=1e-9)


@pytest.mark.parametrize("bp", [{}, {"broad_phase_launches": 1}], ids=["single_launch", "separate_launches"])
def test_ensemble_equals_separate_drums(L, S, bp):
    sc = small_ensemble(S)
    ens = sc.load(L.Sim({**sc.cfg, **bp}))
    ens.step(300)
    pe, we = ens.get_particles(), ens.get_walls()
    drum = sc.meta["drum_of"]
    for k in range(len(COUNTS)):
        one = S.scene_drum_single(sc, k)
        sim = one.load(L.Sim({**one.cfg, **bp}))
        sim.step(300)
        p = sim.get_particles()
        for key in KEYS:
            assert np.array_equal(pe[key][drum == k], p[key]), f"drum {k}: {key} differs"
        assert np.array_equal(we["quat"][k], sim.get_walls()["quat"][0])


def test_ensemble_graph_replays(L, S):
    """Four replays of a 50-step captured graph of the whole ensemble, with eager host reads between
    them, each against 50 eager steps of a second copy."""
    sc = small_ensemble(S)
    eager = sc.load(L.Sim(sc.cfg))
    graph = sc.load(L.Sim(sc.cfg))
    for r in range(4):
        eager.step(50)
        graph.step_graph(50)
        a, b = graph.get_particles(), eager.get_particles()
        for key in KEYS:
            assert np.array_equal(a[key], b[key]), f"replay {r}: {key} differs"
        assert np.array_equal(graph.get_walls()["quat"], eager.get_walls()["quat"])
