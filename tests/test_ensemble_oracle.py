"""Drum ensembles on the CPU: the scene keeps drums apart and carries each drum's own parameters, and
in the oracle one simulation holding all drums evolves every drum exactly as that drum alone."""
import numpy as np

from ensemble_scenes import COUNTS, ES, MUS, RPMS, small_ensemble
from oracle_binding import oracle_sim


def test_scene_layout(S):
    sc = small_ensemble(S)
    drum = sc.meta["drum_of"]
    assert np.array_equal(np.bincount(drum), COUNTS)
    P, W = sc.particles, sc.walls
    assert np.array_equal(P["mu"], np.asarray(MUS)[drum]) and np.array_equal(P["e"], np.asarray(ES)[drum])
    assert np.allclose(W["angvel"][:, 1], np.asarray(RPMS) * 2 * np.pi / 60) and np.array_equal(W["mu"], MUS)
    # every particle inside its own drum's cylinder, far from every other drum's wall grid box
    wall = sc.shapes[W["shape"][0]]
    half = 0.5 * (np.asarray(wall["sdf"].shape) - 1) * wall["spacing"]
    lo, hi = np.asarray(sc.cfg["domain_min"]), np.asarray(sc.cfg["domain_max"])
    for k, c in enumerate(sc.meta["centers"]):
        rel = P["pos"] - c
        mine = drum == k
        assert (np.hypot(rel[mine, 0], rel[mine, 2]) < 0.025).all()
        far = np.maximum(np.abs(rel[~mine, 0]) - half[0], np.abs(rel[~mine, 2]) - half[2])
        assert (far > 5e-3).all()
        assert (c - half >= lo).all() and (c + half <= hi).all()


def test_ensemble_equals_separate_drums(S):
    """60 oracle steps of the five-drum ensemble against each drum simulated on its own, bit for bit."""
    sc = small_ensemble(S)
    ens = sc.load(oracle_sim(sc.cfg))
    wall_contacts = 0
    for _ in range(60):
        ens.step(1)
        wall_contacts += len(ens.contacts(1)["master"])
    assert wall_contacts > 0
    pe = ens.get_particles()
    we = ens.get_walls()
    drum = sc.meta["drum_of"]
    for k in range(len(COUNTS)):
        one = S.scene_drum_single(sc, k)
        sim = one.load(oracle_sim(one.cfg))
        sim.step(60)
        p = sim.get_particles()
        for key in ("pos", "quat", "vel", "angvel", "force", "torque"):
            assert np.array_equal(pe[key][drum == k], p[key]), f"drum {k}: {key}"
        assert np.array_equal(we["quat"][k], sim.get_walls()["quat"][0])
