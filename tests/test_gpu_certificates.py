"""Separation certificates: a candidate pair whose nodes were all culled with some slack is skipped
until the motion bound of its nodes since the pair list was built could have used that slack up.
A skipped pair must never hide a contact.

* spinning, drifting particles against the oracle, step by step (the rotation terms of the bound
  decide when a certificate expires);
* particles resting on a floor that tilts and rises (a wall's rotation and translation);
* a periodic packing run on the GPU with and without certificates: contact lists, forces and the
  whole state stay bit-identical step after step, while most pairs are skipped.
"""
import numpy as np
import pytest

from oracle_binding import oracle_sim
from parity import LISTS, step_and_compare

pytestmark = pytest.mark.gpu


def _pair(sc, L):
    return sc.load(L.Sim(sc.cfg)), sc.load(oracle_sim(sc.cfg))


def test_spinning_particles(L, S):
    sc = S.scene_c1(n=27, lattice=9.8e-3, jitter=0.3e-3, walls=False)
    sc.cfg["gravity"] = (0.0, 0.0, 0.0)
    rng = np.random.default_rng(3)
    sc.particles["angvel"][:] = rng.normal(0.0, 400.0, (sc.n_particles, 3))
    sc.particles["vel"][:] = rng.normal(0.0, 0.05, (sc.n_particles, 3))
    sc.cfg["skin"] = 2e-3
    gpu, ref = _pair(sc, L)
    gpu.profile_begin(80)
    counts, _ = step_and_compare(gpu, ref, 80, lists=(0,))
    prof = gpu.profile_end()
    print("pairs", prof["pairs"], "evaluated", prof["evaluated_pairs"], "rebuilds", prof["rebuilds"])
    assert sum(c["pp"] for c in counts) > 0
    assert prof["evaluated_pairs"] < prof["pairs"]  # some certificates held


def test_tilting_floor(L, S):
    # particles just above the floor (lowest points 0.1-2.1 mm up); the floor rises at 1 m/s while
    # it tilts about its centre, so the pairs' certificates run out one after another
    sc = S.scene_c1(n=36, lattice=10.5e-3, jitter=0.2e-3, box=(0.075, 0.075, 0.3))
    sc.particles["pos"][:, 2] = 5.1e-3
    sc.walls["angvel"][0] = (10.0, -6.0, 1.5)
    sc.walls["vel"][0] = (0.0, 0.0, 1.0)
    sc.cfg["skin"] = 2e-3
    gpu, ref = _pair(sc, L)
    gpu.profile_begin(80)
    # the two sides sum forces in different orders, so states drift apart by rounding over the
    # steps; contact sets stay identical, overlaps agree to 1e-9
    counts, _ = step_and_compare(gpu, ref, 80, lists=(0, 1), ctol=1e-9)
    prof = gpu.profile_end()
    print("pairs", prof["pairs"], "evaluated", prof["evaluated_pairs"], "rebuilds", prof["rebuilds"])
    assert sum(c["pw"] for c in counts) > 10
    assert prof["evaluated_pairs"] < prof["pairs"]


def _state_equal(a: dict, b: dict, what: str):
    for k in ("pos", "quat", "vel", "angvel", "force", "torque"):
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


def test_same_as_full_evaluation(L, S):
    # 3000 particles settling in a periodic cell with a floor; certificates on and off
    sc = S.scene_c2(n=3000, n_types=4, n_grid=32, n_nodes=256, cell_xy=0.036, gap=-0.12, seed=9)
    S.remove_initial_overlaps(lambda c: L.Sim(c), sc, log=lambda m: None)
    on = sc.load(L.Sim(sc.cfg))
    off = sc.load(L.Sim({**sc.cfg, "no_certificates": 1}))
    on.profile_begin(300)
    total_contacts = 0
    for block in range(6):
        on.step(50)
        off.step(50)
        for l in range(5):
            ca, cb = on.contacts(l), off.contacts(l)
            for k in ca:
                assert np.array_equal(ca[k], cb[k]), f"step {50 * block + 49} list {LISTS[l]}: {k} differs"
            total_contacts += len(ca["master"])
        _state_equal(on.get_particles(), off.get_particles(), f"step {50 * block + 49}")
    prof = on.profile_end()
    frac = prof["evaluated_pairs"] / max(prof["pairs"], 1)
    print("pairs", prof["pairs"], "evaluated", prof["evaluated_pairs"], f"({frac:.3f})", "rebuilds",
          prof["rebuilds"], "contacts at checkpoints", total_contacts)
    assert total_contacts > 0
    assert frac < 0.5


def test_certificates_through_rebuilds(L, S):
    # A rebuild every ~10 steps (small skin, fast lateral motion, periodic wraps): particle pairs of
    # each new list inherit what is left of their certificates.  Against full evaluation every
    # step, contacts and states stay bit-identical.
    sc = S.scene_c2(n=1500, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.05, gap=-0.1, seed=17)
    rng = np.random.default_rng(21)
    sc.particles["vel"][:, :2] = rng.normal(0.0, 2.0, (sc.n_particles, 2))
    sc.particles["angvel"][:] = rng.normal(0.0, 50.0, (sc.n_particles, 3))
    sc.cfg["skin"] = 2e-4
    cert = sc.load(L.Sim(sc.cfg))
    full = sc.load(L.Sim({**sc.cfg, "no_certificates": 1}))
    cert.profile_begin(150)
    full.profile_begin(150)
    for block in range(15):
        cert.step(10)
        full.step(10)
        for l in range(5):
            ca, cb = cert.contacts(l), full.contacts(l)
            for k in ca:
                assert np.array_equal(ca[k], cb[k]), f"step {10 * block + 9} list {LISTS[l]}: {k} differs"
        _state_equal(cert.get_particles(), full.get_particles(), f"step {10 * block + 9}")
    pc, pf = cert.profile_end(), full.profile_end()
    print("rebuilds", pc["rebuilds"], "pairs", pc["pairs"], "evaluated", pc["evaluated_pairs"])
    assert pc["rebuilds"] == pf["rebuilds"] and pc["rebuilds"] >= 8
    assert pc["evaluated_pairs"] < 0.5 * pc["pairs"]
