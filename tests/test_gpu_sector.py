"""Quarter-sector rotational periodicity on the GPU against the oracle (run on an MI355X).

Superquadrics in a square quarter-sector cell with a floor; bodies are sent across the x = min and
y = min faces, where they turn by -90 and +90 degrees about the corner's z axis (position, velocity
and orientation), and contacts with the three turned image sets (+90, 180, -90 degrees; lists GX, GY,
GXY) are compared with the oracle's step by step: contact sets bit for bit, forces and torques within
1e-5 relative."""
import numpy as np
import pytest

from oracle_binding import oracle_sim
from parity import step_and_compare

pytestmark = pytest.mark.gpu


def _sector_scene(S, seed=17):
    sc = S.scene_c2(n=500, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.05, gap=-0.15, seed=seed)
    sc.cfg["periodic_x"] = 0
    sc.cfg["periodic_y"] = 0
    sc.cfg["periodic_sector"] = 1
    sc.cfg["gravity"] = (0.0, 0.0, 0.0)
    P = sc.particles
    rng = np.random.default_rng(seed)
    P["vel"][:, :2] = rng.normal(0.0, 2.0, (sc.n_particles, 2))
    lo = np.array(sc.cfg["domain_min"])
    near_x = P["pos"][:, 0] - lo[0] < 6e-3
    near_y = P["pos"][:, 1] - lo[1] < 6e-3
    P["vel"][near_x, 0] = -25.0  # across x = min within a few steps
    P["vel"][near_y & ~near_x, 1] = -25.0
    return sc


def test_sector_matches_oracle(L, S):
    sc = _sector_scene(S)
    gpu = sc.load(L.Sim(sc.cfg))
    ref = sc.load(oracle_sim(sc.cfg))
    q0 = ref.get_particles()["quat"].copy()
    counts, worst = step_and_compare(gpu, ref, 40, lists=(0, 1, 2, 3, 4), log=print)
    ghost = [c["gx"] + c["gy"] + c["gxy"] for c in counts]
    print("image contacts per step", ghost, "worst", worst)
    assert sum(ghost) > 0
    # bodies turned by a quarter: their orientation changed by far more than their spin allows
    q1 = ref.get_particles()["quat"]
    turned = np.abs(np.abs(np.sum(q0 * q1, axis=1)) - 1.0) > 0.05
    assert turned.sum() > 0
    assert np.array_equal(gpu.get_particles()["quat"], q1) or np.abs(gpu.get_particles()["quat"] - q1).max() < 1e-12


def test_sector_rejects_xy(L, S):
    sc = _sector_scene(S)
    sc.cfg["periodic_x"] = 1
    with pytest.raises(L.LSDEMError):
        L.Sim(sc.cfg)
