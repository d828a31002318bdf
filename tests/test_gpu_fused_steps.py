"""Consecutive steps of one call share a launch at each step boundary: the closing sums and second
half-kick of one step run with the first half-kick and drift of the next.  Against separate
launches (`no_fused_steps`), states, contacts and counters are identical after every call, through
rebuilds, periodic wraps, a moving and tilting wall, and poses set between calls."""
import numpy as np
import pytest

from parity import LISTS

pytestmark = pytest.mark.gpu

CHUNKS = (1, 2, 7, 13, 30, 1, 25, 3)


def _same(a, b, where):
    assert a.counters() == b.counters(), f"{where}: counters differ"
    for l in range(5):
        ca, cb = a.contacts(l), b.contacts(l)
        for k in ca:
            assert np.array_equal(ca[k], cb[k]), f"{where} list {LISTS[l]}: {k} differs"
    pa, pb = a.get_particles(), b.get_particles()
    for k in ("pos", "quat", "vel", "angvel", "force", "torque"):
        assert np.array_equal(pa[k], pb[k]), f"{where}: {k} differs"
    wa, wb = a.get_walls(), b.get_walls()
    for k in wa:
        assert np.array_equal(wa[k], wb[k]), f"{where}: wall {k} differs"


def test_fused_periodic_packing(L, S):
    sc = S.scene_c2(n=1500, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.05, gap=-0.1, seed=17)
    rng = np.random.default_rng(9)
    sc.particles["vel"][:, :2] = rng.normal(0.0, 5.0, (sc.n_particles, 2))  # frequent rebuilds and wraps
    sc.cfg["skin"] = 1e-3
    fu = sc.load(L.Sim(sc.cfg))
    se = sc.load(L.Sim({**sc.cfg, "no_fused_steps": 1}))
    fu.profile_begin(sum(CHUNKS))
    for i, n in enumerate(CHUNKS):
        fu.step(n)
        se.step(n)
        _same(fu, se, f"call {i} ({n} steps)")
    prof = fu.profile_end()
    print("rebuilds", prof["rebuilds"], "evaluated", prof["evaluated_pairs"], "steps", prof["steps"])
    assert prof["rebuilds"] >= 5
    assert sum(fu.counters().values()) > 0


def test_fused_moving_wall_and_poses_set_between_calls(L, S):
    sc = S.scene_c1(n=36, lattice=10.5e-3, jitter=0.2e-3, box=(0.075, 0.075, 0.3))
    sc.particles["pos"][:, 2] = 5.1e-3
    sc.walls["angvel"][0] = (10.0, -6.0, 1.5)
    sc.walls["vel"][0] = (0.0, 0.0, 1.0)
    sc.cfg["skin"] = 2e-3
    fu = sc.load(L.Sim(sc.cfg))
    se = sc.load(L.Sim({**sc.cfg, "no_fused_steps": 1}))
    pw = 0
    for i, n in enumerate(CHUNKS):
        if i == 4:
            # poses changed between calls: the next step is a forced rebuild
            for s in (fu, se):
                p = s.get_particles()
                pos = p["pos"].copy()
                pos[0, 2] += 1e-3
                s.set_particles(pos=pos)
        fu.step(n)
        se.step(n)
        _same(fu, se, f"call {i} ({n} steps)")
        pw += len(fu.contacts(1)["master"])
    assert pw > 0


def test_fused_after_hook_and_forces_set_between_calls(L, S):
    # a damping hook on some calls (every step then launches separately and the hook adds to the
    # stored sums), off on others; forces and torques set from the host between calls
    import ctypes as C
    sc = S.scene_c1(n=27, lattice=9.8e-3, jitter=0.2e-3)
    fu = sc.load(L.Sim(sc.cfg))
    se = sc.load(L.Sim({**sc.cfg, "no_fused_steps": 1}))
    f = fu.lib.launchAddGlobalDampingForceTorque
    f.restype = None
    f.argtypes = [C.c_void_p] * 4 + [C.c_double, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p]

    def hook(time, v):
        f(v.force, v.torque, v.velocity, v.angular_velocity, 0.3, v.num, v.grid_dim, v.block_dim, v.stream)
    rng = np.random.default_rng(4)
    nonzero = 0
    for i, n in enumerate(CHUNKS):
        for s in (fu, se):
            s.set_external_force(hook if i in (1, 2, 5) else None)
        if i == 3:
            force = rng.normal(0.0, 1e-3, (sc.n_particles, 3))
            torque = rng.normal(0.0, 1e-6, (sc.n_particles, 3))
            for s in (fu, se):
                s.set_particles(force=force, torque=torque)
        fu.step(n)
        se.step(n)
        _same(fu, se, f"call {i} ({n} steps)")
        nonzero += int(np.count_nonzero(fu.get_particles()["force"]))
    assert nonzero > 0


def test_block_size_not_a_wavefront_multiple(L, S):
    # the body kernels round the app's block size up to whole wavefronts; results do not change
    sc = S.scene_c1(n=36, lattice=10.5e-3, jitter=0.2e-3, box=(0.075, 0.075, 0.3))
    sc.particles["pos"][:, 2] = 5.1e-3
    a = sc.load(L.Sim({**sc.cfg, "block_size": 100}))
    b = sc.load(L.Sim(sc.cfg))
    for i, n in enumerate((1, 20, 40)):
        a.step(n)
        b.step(n)
        _same(a, b, f"call {i} ({n} steps)")


def test_sampled_profile(L, S):
    # phases timed on every third step; counters over all of them
    sc = S.scene_c1(n=27, lattice=9.8e-3, jitter=0.2e-3)
    sim = sc.load(L.Sim(sc.cfg))
    sim.profile_begin(10, every=3)
    sim.step(30)
    p = sim.profile_end()
    assert p["steps"] == 30 and p["sampled_steps"] == 10
    assert all(v >= 0.0 for v in p["ms"].values()) and sum(p["ms"].values()) > 0.0
    sim.profile_begin(40)
    sim.step(30)
    q = sim.profile_end()
    assert q["steps"] == 30 and q["sampled_steps"] == 30


@pytest.mark.parametrize("launches", [0, 1])
def test_graph_steps_match_eager(L, S, launches):
    # the same calls replayed from a captured graph, with the program's own reads between the
    # replays; every replay against eager steps (single-launch and separate-launch broad phase)
    sc = S.scene_c2(n=600, n_types=3, n_grid=24, n_nodes=128, cell_xy=0.05, gap=-0.1, seed=23)
    rng = np.random.default_rng(2)
    sc.particles["vel"][:, :2] = rng.normal(0.0, 2.0, (sc.n_particles, 2))
    sc.cfg["skin"] = 5e-4
    cfg = {**sc.cfg, "broad_phase_launches": launches}
    gr = sc.load(L.Sim(cfg))
    ea = sc.load(L.Sim(cfg))
    gr.step(2)  # the first steps build the pairs (a pending rebuild request runs eagerly)
    ea.step(2)
    for i in range(6):
        gr.step_graph(20)
        ea.step(20)
        _same(gr, ea, f"replay {i}")
        assert gr.time() == ea.time()
