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
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 5, lists=(0, 1), log=print)
    assert sum(c["pp"] for c in counts) > 100
    print("worst", worst)


def test_c1_floor_contacts(L, S):
    # particles resting on the floor and side walls: particle-wall lists exercised
    sc = S.scene_c1(n=49, lattice=10.5e-3, jitter=0.2e-3, box=(0.075, 0.075, 0.3))
    sc.particles["pos"][:, 2] = 2.6e-3
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 4, lists=(0, 1), log=print)
    assert sum(c["pw"] for c in counts) > 10


def test_c2_periodic_small(L, S):
    # periodic x/y with a floor: ghost lists +x, +y, +x+y
    sc = S.scene_c2(n=600, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.06, gap=-0.15)
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 4, lists=(0, 1, 2, 3, 4), log=print)
    assert sum(c["gx"] + c["gy"] for c in counts) > 0


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
