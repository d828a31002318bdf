"""GPU library vs the oracle, from identical initial states (run on an MI355X).

Contact sets must match bit for bit after sorting by (master node, slave); overlap, point and
normal agree to 1e-12 relative (same FP64 expressions, correctly rounded sqrt on both sides);
per-step forces and torques agree to 1e-5 relative to the step's largest body force (the
two sides sum contributions in different orders).
"""
import numpy as np
import pytest

from oracle_binding import oracle_sim
from parity import step_and_compare

pytestmark = pytest.mark.gpu


def _pair(sc, L):
    gpu = sc.load(L.Sim(sc.cfg))
    ref = sc.load(oracle_sim(sc.cfg))
    return gpu, ref


def test_c1_compressed_walls(L, S):
    # 64 superellipsoids on a lattice tighter than their bounding spheres: many contacts at step 1
    sc = S.scene_c1(n=64, lattice=9.5e-3, jitter=0.3e-3)
    sc.cfg["skin"] = 1e-3  # candidate pairs reused across the steps
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 12, lists=(0, 1), log=print)
    assert sum(c["pp"] for c in counts) > 100
    print("worst", worst)


def test_c1_floor_contacts(L, S):
    # particles resting on the floor and side walls: particle-wall lists exercised
    sc = S.scene_c1(n=49, lattice=10.5e-3, jitter=0.2e-3, box=(0.075, 0.075, 0.3))
    sc.particles["pos"][:, 2] = 2.6e-3
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 4, lists=(0, 1), log=print)
    assert sum(c["pw"] for c in counts) > 10


@pytest.mark.parametrize("skin", [0.0, -1.0, 2e-3])
def test_c2_periodic_small(L, S, skin):
    # periodic x/y with a floor: ghost lists +x, +y, +x+y.  skin 0 rebuilds the candidate pairs
    # every step, -1 is the default, 2 mm (a third of a radius) keeps them for many steps
    sc = S.scene_c2(n=600, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.06, gap=-0.15)
    sc.cfg["skin"] = skin
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 4 if skin == 0.0 else 25, lists=(0, 1, 2, 3, 4), log=print)
    assert sum(c["gx"] + c["gy"] for c in counts) > 0


def test_reused_pairs_through_wraps(L, S):
    # fast particles crossing the periodic faces while the candidate pairs are reused (large skin)
    sc = S.scene_c2(n=300, n_types=2, n_grid=24, n_nodes=128, cell_xy=0.05, gap=-0.1)
    rng = np.random.default_rng(11)
    sc.particles["vel"][:, :2] = rng.normal(0.0, 20.0, (sc.n_particles, 2))
    sc.cfg["skin"] = 3e-3
    gpu, ref = _pair(sc, L)
    x0 = ref.get_particles()["pos"].copy()
    step_and_compare(gpu, ref, 40, lists=(0, 1, 2, 3, 4))
    x1 = ref.get_particles()["pos"]
    assert np.any(np.abs(x1[:, 0] - x0[:, 0]) > 0.02)  # some particles wrapped around


def test_history_persists(L, S):
    # friction springs carried across 20 steps must stay equal
    sc = S.scene_c1(n=27, lattice=9.8e-3, jitter=0.2e-3, walls=False)
    sc.cfg["gravity"] = (0.0, 0.0, 0.0)
    gpu, ref = _pair(sc, L)
    step_and_compare(gpu, ref, 20, lists=(0,))
    a = gpu.contacts(0)
    assert np.abs(a["spring"]).max() > 0


def test_c3_bonded_compression(L, S):
    # bonded agglomerate squeezed between closing walls: bond forces every step and every break
    # event on the same step as the oracle
    sc = S.scene_c3(n=27, seed=5, exponents=(2.0,), r_range=(2.95e-3, 3e-3), pitch=6.02e-3, jitter=0.0,
                    wall_speed=0.5, weibull_scale=2e4, n_grid=24, n_nodes=128)
    assert len(sc.bonds) > 40
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 120, lists=(0, 1), log=print, bonds=True)
    broken = [c["broken"] for c in counts]
    assert broken[0] < broken[-1] and broken[-1] >= 40
