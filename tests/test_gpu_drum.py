"""A rotating drum (the CylinderWall of the reference's drum apps, level set positive inside) holding
superellipsoids: contact sets against the oracle step by step while the drum turns, and a few drums
stepped concurrently from several host threads giving the same states as one after another."""
import threading

import numpy as np
import pytest

from oracle_binding import oracle_sim
from parity import step_and_compare

pytestmark = pytest.mark.gpu


def _small_drum(S, seed=3, rpm=60.0, mu=0.5, e=0.5):
    sc = S.scene_drum(n=40, radius=0.025, depth=0.016, n_grid=20, n_nodes=96, wall_spacing=5e-4, rpm=rpm, seed=seed,
                      semi_axes=(1.5e-3, 2.5e-3), exponents=(2.0, 4.0), mu=mu, restitution=e)
    # three particles, unrotated, rest on the bottom of the mantle, 20 um into it
    P = sc.particles
    for k, x in enumerate((-6e-3, 0.0, 6e-3)):
        ext = -sc.shapes[P["shape"][k]]["nodes"][:, 2].min()  # lowest node below the centre
        P["quat"][k] = (1.0, 0.0, 0.0, 0.0)
        P["pos"][k] = (x, 0.0, -np.sqrt((0.025 - ext + 2e-5) ** 2 - x * x))
    return sc


def test_drum_matches_oracle(L, S):
    sc = _small_drum(S)
    gpu = sc.load(L.Sim(sc.cfg))
    ref = sc.load(oracle_sim(sc.cfg))
    counts, _ = step_and_compare(gpu, ref, 40, lists=(0, 1), ctol=1e-9)
    assert sum(c["pw"] for c in counts) > 0
    wa, wb = gpu.get_walls(), ref.get_walls()
    assert np.array_equal(wa["quat"], wb["quat"])  # the drum turned the same way


@pytest.mark.parametrize("bp", [{"broad_phase_launches": 1}, {"broad_phase_launches": 0, "broad_phase_blocks": 8}],
                         ids=["separate_launches", "single_launch_8_blocks"])
def test_drums_concurrently(L, S, bp):
    cfgs = [dict(seed=11 + k, rpm=10.0 + 15.0 * k, mu=0.2 + 0.2 * k, e=0.3 + 0.2 * k) for k in range(4)]
    scs = [_small_drum(S, **c) for c in cfgs]
    # separate broad-phase launches (concurrent contexts never wait on each other's grid barriers), or
    # single launches of 8 blocks (the four contexts' blocks fit on the GPU together)
    conc = [sc.load(L.Sim({**sc.cfg, **bp})) for sc in scs]
    seq = [sc.load(L.Sim({**sc.cfg, **bp})) for sc in scs]
    errs = []

    def run(sim):
        try:
            sim.step(200)
        except Exception as ex:  # reported below
            errs.append(ex)
    th = [threading.Thread(target=run, args=(s,)) for s in conc]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for s in seq:
        s.step(200)
    for k, (a, b) in enumerate(zip(conc, seq)):
        pa, pb = a.get_particles(), b.get_particles()
        for key in ("pos", "quat", "vel", "angvel", "force", "torque"):
            assert np.array_equal(pa[key], pb[key]), f"drum {k}: {key} differs"
