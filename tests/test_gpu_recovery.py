"""Error paths of a call, forced from the host with every kernel still exiting normally.

A call whose pairs or candidate slots do not fit, or whose broad-phase grid barrier gives up, halts
on the device (every later kernel returns at entry); the library restores the call's start, grows
the buffers (or switches to separate broad-phase launches) and runs the call again.  The results
must be those of a context that had the room from the start, bit for bit.  Also: the broad phase
under hipLaunchCooperativeKernel, and the pre-loop build against the oracle.
"""
import numpy as np
import pytest

from oracle_binding import oracle_sim
from parity import LISTS, step_and_compare

pytestmark = pytest.mark.gpu


def _scene(S):
    sc = S.scene_c2(n=600, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.06, gap=-0.15, seed=23)
    rng = np.random.default_rng(2)
    sc.particles["vel"][:, :2] = rng.normal(0.0, 3.0, (sc.n_particles, 2))  # rebuilds and wraps
    sc.cfg["skin"] = 1e-3
    return sc


def _same(a, b, where):
    pa, pb = a.get_particles(), b.get_particles()
    for k in ("pos", "quat", "vel", "angvel", "force", "torque"):
        assert np.array_equal(pa[k], pb[k]), f"{where}: {k} differs"
    for l in range(5):
        ca, cb = a.contacts(l), b.contacts(l)
        for k in ca:
            assert np.array_equal(ca[k], cb[k]), f"{where} list {LISTS[l]}: {k} differs"


@pytest.mark.parametrize("graph", [False, True])
def test_capacity_recovery_is_exact(L, S, graph):
    sc = _scene(S)
    ref = sc.load(L.Sim(sc.cfg))
    # about 2,600 candidate pairs and 200 candidate slots per step: both capacities overflow
    small = sc.load(L.Sim({**sc.cfg, "pair_capacity": 400, "contact_capacity": 100}))
    c0 = small.counters()
    for call in range(6):
        ref.step(10)
        if graph and call > 0:
            small.step_graph(10)
        else:
            small.step(10)
        print(f"call {call}: counters {small.counters()}", flush=True)
        _same(small, ref, f"call {call}")
    c1 = small.counters()
    assert c1["recoveries"] >= 2
    assert c1["pair_capacity"] > c0["pair_capacity"] and c1["contact_capacity"] > c0["contact_capacity"]
    assert c1["steps"] == ref.counters()["steps"] == 60


def test_barrier_give_up_recovers_with_separate_launches(L, S):
    """barrier_spin = -1: every grid barrier gives up at once, so the first rebuild halts the call;
    it is run again with the separate-launch broad phase."""
    sc = _scene(S)
    ref = sc.load(L.Sim(sc.cfg))
    hurry = sc.load(L.Sim({**sc.cfg, "barrier_spin": -1}))
    for call in range(3):
        ref.step(8)
        hurry.step(8)
        _same(hurry, ref, f"call {call}")
    assert hurry.counters()["recoveries"] >= 1


def test_no_recovery_reports_the_overflow(L, S):
    sc = _scene(S)
    sim = sc.load(L.Sim({**sc.cfg, "pair_capacity": 100, "no_recovery": 1}))
    with pytest.raises(L.LSDEMError) as e:
        sim.step(3)
    assert e.value.code == L.LSDEM_ERR_CAPACITY and "pair" in str(e.value)
    # the same capacity error again (nothing else breaks, nothing hangs)
    with pytest.raises(L.LSDEMError):
        sim.step(1)


def test_cooperative_launch_identical(L, S):
    sc = _scene(S)
    ref = sc.load(L.Sim(sc.cfg))
    coop = sc.load(L.Sim({**sc.cfg, "broad_phase_coop": 1}))
    for call in range(4):
        ref.step(10)
        coop.step(10)
        _same(coop, ref, f"call {call}")
    assert coop.counters()["recoveries"] == 0


def test_preloop_build_matches_oracle(L, S):
    """Particles starting one period outside the cell are wrapped by the pre-loop build; the GPU
    follows the oracle step by step, and begin_phase() alone lists the starting contacts."""
    sc = S.scene_c2(n=400, n_types=2, n_grid=24, n_nodes=128, cell_xy=0.05, gap=-0.12, seed=29)
    lx = sc.cfg["domain_max"][0] - sc.cfg["domain_min"][0]
    ly = sc.cfg["domain_max"][1] - sc.cfg["domain_min"][1]
    sc.particles["pos"][::7, 0] -= lx
    sc.particles["pos"][3::11, 1] += ly
    gpu = sc.load(L.Sim(sc.cfg))
    ref = sc.load(oracle_sim(sc.cfg))
    gpu.begin_phase()
    ref.begin_phase()
    for l in range(5):
        a, b = gpu.contacts(l), ref.contacts(l)
        assert np.array_equal(a["master"], b["master"]) and np.array_equal(a["slave"], b["slave"])
    pa, pb = gpu.get_particles(), ref.get_particles()
    assert np.array_equal(pa["pos"], pb["pos"])  # wrapped, nothing integrated
    counts, worst = step_and_compare(gpu, ref, 15, lists=(0, 1, 2, 3, 4), log=print)
    assert sum(c["pp"] for c in counts) > 0
