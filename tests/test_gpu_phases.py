"""The per-phase entry points (lsdem_phase_*) on the GPU: a step driven phase by phase equals
lsdem_step bit for bit; the cell table of the hash phase equals a brute-force binning of the bodies'
centres; the candidate pairs of the pairs phase hold every body pair whose bounding spheres come
within the skin and no pair beyond it; calls out of order are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "quat", "vel", "angvel", "force", "torque")


def _periodic(S):
    return S.scene_c2(n=500, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.05, gap=-0.15, seed=31)


def _bonded(S):
    return S.scene_c3(n=27, seed=5, exponents=(2.0,), r_range=(2.95e-3, 3e-3), pitch=6.02e-3, jitter=0.0,
                      wall_speed=0.5, weibull_scale=2e4, n_grid=24, n_nodes=128)


def _phased_step(sim):
    for name in sim.STEP_PHASES:
        sim.phase(name)


@pytest.mark.parametrize("make", [_periodic, _bonded], ids=["periodic_floor", "bonded_walls"])
def test_phases_equal_step(L, S, make):
    sc = make(S)
    a = sc.load(L.Sim(sc.cfg))
    b = sc.load(L.Sim(sc.cfg))
    contacts = 0
    for k in range(40):
        _phased_step(a)
        b.step(1)
        pa, pb = a.get_particles(), b.get_particles()
        for key in KEYS:
            assert np.array_equal(pa[key], pb[key]), f"step {k}: {key} differs"
        for lst in (0, 1, 2, 3, 4):
            ca, cb = a.contacts(lst), b.contacts(lst)
            assert np.array_equal(ca["master"], cb["master"]) and np.array_equal(ca["slave"], cb["slave"])
            assert np.array_equal(ca["spring"], cb["spring"])
            contacts += len(ca["master"])
        if sc.bonds:
            assert np.array_equal(a.bonds()["active"], b.bonds()["active"])
    assert contacts > 0
    assert a.time() == b.time()


def test_hash_cell_table_brute_force(L, S):
    sc = _periodic(S)
    sim = sc.load(L.Sim(sc.cfg))
    sim.step(5)
    for k in range(3):
        sim.phase("integrate")
        sim.phase("hash")
        t = sim.cell_table()
        pos = sim.get_particles()["pos"]
        dims = t["dims"]
        ncell = int(np.prod(dims))
        # the cell of each wrapped centre, clamped to the grid, linear index x fastest
        c = np.floor((pos - t["origin"]) * (1.0 / t["cell"]))
        c = np.minimum(np.maximum(c, 0.0), dims - 1).astype(np.int64)
        lin = c[:, 0] + dims[0] * (c[:, 1] + dims[1] * c[:, 2])
        assert np.array_equal(t["cell_of"], lin), f"call {k}: cells differ"
        assert np.array_equal(t["sorted"], np.lexsort((np.arange(len(lin)), lin))), f"call {k}: order differs"
        start = np.concatenate([[0], np.cumsum(np.bincount(lin, minlength=ncell))])
        assert np.array_equal(t["cell_start"], start), f"call {k}: cell table differs"
        assert len(np.unique(lin)) > 50  # bodies spread over many cells
        for name in sim.STEP_PHASES[2:]:
            sim.phase(name)


def test_pairs_brute_force(L, S):
    sc = _periodic(S)
    sc.cfg["skin"] = 1.5e-3
    sc.cfg["periodic_x"] = 0
    sc.cfg["periodic_y"] = 0
    sim = sc.load(L.Sim(sc.cfg))
    sim.step(3)
    sim.phase("integrate")
    sim.phase("hash")
    sim.phase("pairs")
    pr = sim.pairs()
    pos = sim.get_particles()["pos"]
    P = sc.particles
    rad = np.array([np.linalg.norm(sc.shapes[s]["nodes"] * P["scale"][i], axis=1).max()
                    for i, s in enumerate(P["shape"])])
    lo, hi = np.asarray(sc.cfg["domain_min"]), np.asarray(sc.cfg["domain_max"])
    inside = ((pos >= lo) & (pos < hi)).all(axis=1)
    assert inside.all()  # (no body outside the domain, so every body is a possible partner)
    pp = pr["type"] == 0
    got = set(zip(pr["master"][pp].tolist(), pr["partner"][pp].tolist()))
    assert len(got) == pp.sum()  # no duplicates
    assert all(a < b for a, b in got)
    # masters in ascending order
    assert np.all(np.diff(pr["master"]) >= 0)
    skin = sc.cfg["skin"]
    i, j = np.triu_indices(len(pos), 1)
    d = np.linalg.norm(pos[i] - pos[j], axis=1)
    reach = rad[i] + rad[j] + skin
    # (the search tests FP32 centres with a margin of a few 1e-7 m)
    inner, outer = d <= reach - 1e-6, d <= reach * (1 + 1e-5) + 1e-6
    must = set(zip(i[inner].tolist(), j[inner].tolist()))
    may = set(zip(i[outer].tolist(), j[outer].tolist()))
    assert must <= got, f"{len(must - got)} pairs within the skin are missing"
    assert got <= may, f"{len(got - may)} pairs beyond the skin"
    assert len(must) > 200
    for name in sim.STEP_PHASES[3:]:
        sim.phase(name)


def test_out_of_order_refused(L, S):
    sc = _bonded(S)
    sim = sc.load(L.Sim(sc.cfg))
    with pytest.raises(L.LSDEMError, match="out of order"):
        sim.phase("hash")
    sim.phase("integrate")
    with pytest.raises(L.LSDEMError, match="unfinished"):
        sim.step(1)
    with pytest.raises(L.LSDEMError, match="out of order"):
        sim.phase("narrow")
    for name in sim.STEP_PHASES[1:]:
        sim.phase(name)
    sim.step(2)  # whole steps again once the phased step has ended


def test_wall_device_view(L, S):
    """getWall()'s device arrays: they hold the walls' poses, and a velocity written into them moves
    the wall from the next step on."""
    from hip_rt import Hip
    hip = Hip()
    sc = _bonded(S)
    sim = sc.load(L.Sim(sc.cfg))
    sim.step(10)
    v = sim.wall_device_view()
    w = sim.get_walls()
    nW = len(w["pos"])
    assert v.num == nW and v.block_dim > 0 and v.grid_dim * v.block_dim >= nW
    assert np.array_equal(hip.read(v.position, np.zeros((nW, 3))), w["pos"])
    assert np.array_equal(hip.read(v.quaternion, np.zeros((nW, 4))), w["quat"])
    assert np.array_equal(hip.read(v.velocity, np.zeros((nW, 3))), w["vel"])
    vel = w["vel"].copy()
    vel[0] = (0.0, 0.0, 0.25)
    hip.write(v.velocity, vel)
    sim.step(4)
    w2 = sim.get_walls()
    assert np.array_equal(w2["vel"], vel)
    dt = sc.cfg["dt"]
    assert np.allclose(w2["pos"][0] - w["pos"][0], (0.0, 0.0, 4 * 0.25 * dt), rtol=0, atol=1e-15)
