"""GPU library vs the oracle at the benchmark configurations' own shape and material parameters
(run on an MI355X).  Sizes are cut to what the oracle steps in seconds; every other parameter is
the configuration's:

- 200k-particle roofline workload: 10 superquadric types (exponents U[2,8], axis ratios
  (1, U[.5,1], U[.5,1])), FP32 SDF 50^3, 512 nodes, radii lognormal (median 3 mm, sigma 0.25),
  periodic x/y with a floor, kn 1e5, kt 0.4 kn, e 0.5, mu 0.5, dt 5e-6;
- bonded agglomerate: exponents 2, 3, 6, equivalent radius U[2, 3] mm, SDF 40^3, 400 nodes, bonds
  between surfaces closer than 0.1 mm with radius 0.8 x the smaller radius, E 1e8 Pa, Weibull(5,
  1e6 Pa) strengths, walls closing at 0.05 m/s; the lattice starts compressed so that breaks fall
  inside the test;
- rotating drum: radius 0.1 m, depth 0.05 m, superellipsoids with exponents U[2,6] and semi-axes
  3-5 mm, SDF 32^3, 300 nodes, dt 1e-5, mu and e drawn from the configuration's ranges.

Contact sets bit for bit after sorting, forces and torques within 1e-5 relative, bond break events
on the same step (see parity.py)."""
import numpy as np
import pytest

from oracle_binding import oracle_sim
from parity import step_and_compare

pytestmark = pytest.mark.gpu


def _pair(sc, L, threads=8):
    gpu = sc.load(L.Sim(sc.cfg))
    ref = sc.load(oracle_sim(sc.cfg, threads=threads))
    return gpu, ref


def test_c2_configuration(L, S):
    sc = S.scene_c2(n=3000, cell_xy=0.1, gap=-0.15, seed=31)  # 10 types, 50^3, 512 nodes (defaults)
    assert len(sc.shapes) == 11 and sc.shapes[0]["sdf"].shape == (50, 50, 50) and len(sc.shapes[0]["nodes"]) == 512
    moved = S.remove_initial_overlaps(lambda c: L.Sim(c), sc)
    rng = np.random.default_rng(4)
    sc.particles["vel"] = rng.normal(0.0, 0.3, (sc.n_particles, 3))
    # lower the bed until its lowest boundary node is 10 um into the floor (z = 0): floor contacts
    P = sc.particles
    low = min(float(P["pos"][i, 2] + (S.quat_rotate(P["quat"][i][None, :], sc.shapes[P["shape"][i]]["nodes"])[:, 2]
                                      * P["scale"][i]).min()) for i in range(sc.n_particles))
    P["pos"][:, 2] -= low + 1e-5
    print(f"{sc.n_particles} particles, {moved} relocated, domain {sc.cfg['domain_max']}", flush=True)
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 25, lists=(0, 1, 2, 3, 4), log=print)
    assert sum(c["pp"] for c in counts) > 500
    assert sum(c["gx"] + c["gy"] + c["gxy"] for c in counts) > 0
    assert sum(c["pw"] for c in counts) > 0
    print("worst", worst)


def test_c2_large_body_pass(L, S):
    """More bodies (70,000) than the small-simulation body pass takes (65,536): the large body pass,
    whose step close (counter sums, flags) runs in a block of its own, against the oracle."""
    sc = S.scene_c2(n=70_000, cell_xy=0.2, gap=-0.15, seed=37)
    assert sc.n_particles > 65_536
    moved = S.remove_initial_overlaps(lambda c: L.Sim(c), sc)
    rng = np.random.default_rng(5)
    sc.particles["vel"] = rng.normal(0.0, 0.3, (sc.n_particles, 3))
    # the lowest layer 10 um into the floor: floor contacts
    P = sc.particles
    bottom = np.nonzero(P["pos"][:, 2] < P["pos"][:, 2].min() + 0.01)[0]
    low = min(float(P["pos"][i, 2] + (S.quat_rotate(P["quat"][i][None, :], sc.shapes[P["shape"][i]]["nodes"])[:, 2]
                                      * P["scale"][i]).min()) for i in bottom)
    P["pos"][:, 2] -= low + 1e-5
    print(f"{sc.n_particles} particles, {moved} relocated, domain {sc.cfg['domain_max']}", flush=True)
    gpu, ref = _pair(sc, L, threads=16)
    counts, worst = step_and_compare(gpu, ref, 8, lists=(0, 1, 2, 3, 4), log=print)
    assert sum(c["pp"] for c in counts) > 500
    assert sum(c["pw"] for c in counts) > 0
    print("worst", worst)


def test_c3_configuration(L, S):
    # a lattice tighter than the largest particles (pitch 5.6 mm for radii up to 3 mm): the packing
    # starts compressed, its contacts push the bonded particles apart and bonds break over many steps
    sc = S.scene_c3(n=216, seed=13, pitch=5.6e-3, jitter=0.2, bonds=False)
    assert sc.shapes[0]["sdf"].shape == (40, 40, 40) and len(sc.shapes[0]["nodes"]) == 400
    rng = np.random.default_rng(8)
    sc.bonds = S.bond_pairs_by_gap(sc, 1e-4, 0.8, 1e8, 0.3, rng, 5.0, 1e6)
    print(f"{sc.n_particles} particles, {len(sc.bonds)} bonds", flush=True)
    assert len(sc.bonds) > 60
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 120, lists=(0, 1), log=print, bonds=True)
    broken = [c["broken"] for c in counts]
    print("broken per step", broken[::10], "of", len(sc.bonds))
    assert len(set(broken)) > 8 and broken[-1] > len(sc.bonds) // 3  # breaks spread over many steps


def test_c5_drum_configuration(L, S):
    rng = np.random.default_rng(12)
    sc = S.scene_drum(n=300, seed=21, rpm=float(rng.uniform(1, 30)), mu=float(rng.uniform(0.1, 0.8)),
                      restitution=float(rng.uniform(0.3, 0.9)))
    assert sc.shapes[0]["sdf"].shape == (32, 32, 32) and len(sc.shapes[0]["nodes"]) == 300
    # the lowest particles rest 20 um into the mantle, so the turning wall drags them from step 1
    P = sc.particles
    low = np.argsort(P["pos"][:, 2])[:6]
    for k, i in enumerate(low):
        x = -0.012 + 0.005 * k
        ext = -sc.shapes[P["shape"][i]]["nodes"][:, 2].min()
        P["quat"][i] = (1.0, 0.0, 0.0, 0.0)
        P["pos"][i] = (x, P["pos"][i][1], -np.sqrt((0.1 - ext + 2e-5) ** 2 - x * x))
    gpu, ref = _pair(sc, L)
    counts, worst = step_and_compare(gpu, ref, 40, lists=(0, 1), ctol=1e-9, log=print)
    assert sum(c["pw"] for c in counts) > 0
    wa, wb = gpu.get_walls(), ref.get_walls()
    assert np.array_equal(wa["quat"], wb["quat"])
