"""Pins the oracle (host restatement of the reference LS-DEM step) without running the reference.

The reference ships no tests, fixtures or golden vectors (it is CUDA + Thrust and cannot be
built here), so each oracle function is pinned by mathematics that a plausible mistake breaks:

  function (oracle)                      pinned by
  linearContact / slidingSpring          closed forms: F = kn*delta*n at rest; Coulomb cap |Ft| = mu|Fn|;
                                         restitution and contact duration of a damped linear oscillator
  kick / drift / quaternionIntegration   free flight is exact for constant g; |q| = 1; w constant w/o torque
  interpValue / interpGradient           linear fields reproduced exactly; hand KAT of the `int x` quirk
  node-vs-SDF detection (pp, pw, ghost)  brute-force sets with an independent formulation (rotation
                                         matrices, all pairs, no grid); literal traversal == pair traversal
  keys / sort / cell table               brute-force keys, stable order, start/end by linear scan
  massFromSdf                            sphere mass and inertia vs (4/3) pi r^3 rho and (2/5) m r^2
  forces of the three contact kinds      Newton's third law (pp sums to zero), wall reaction absent
  bonds (vbond)                          axial stiffness B1 = EA/L under small stretch; break threshold
"""
import ctypes as C

import numpy as np
import pytest

from oracle_binding import oracle_lib, oracle_sim

D3 = C.c_double * 3


def sphere_shape(r, n_grid, h, nodes):
    o = -0.5 * (n_grid - 1) * h
    ax = o + h * np.arange(n_grid)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    sdf = (np.sqrt(X * X + Y * Y + Z * Z) - r).astype(np.float32)
    return sdf, np.array([o, o, o]), h, np.asarray(nodes, float)


def axis_nodes(r):
    return r * np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)


def one_particle_scene(S, shape, pos, vel, m, kn, e, mu=0.0, ks=0.0, g=(0, 0, 0), dt=1e-6, walls=True,
                       inv_inertia=None):
    sc = S.Scene(cfg={"domain_min": (-0.05, -0.05, -0.05), "domain_max": (0.05, 0.05, 0.05), "gravity": g,
                      "dt": dt})
    sc.shapes.append({"sdf": shape[0], "origin": shape[1], "spacing": shape[2], "nodes": shape[3]})
    n = len(pos)
    P = S._particles(n)
    P["pos"][:] = pos
    P["vel"][:] = vel
    P["quat"][:] = (1, 0, 0, 0)
    P["inv_mass"][:] = 1.0 / m
    P["inv_inertia"][:] = inv_inertia if inv_inertia is not None else (1.0 / (0.4 * m * 25e-6),) * 3 + (0, 0, 0)
    P["kn"][:] = kn
    P["ks"][:] = ks
    P["mu"][:] = mu
    P["e"][:] = e
    sc.particles = P
    sc.walls = S.floor_wall(sc.shapes, (0, 0), 0.0, 0.04, mu, 1.0, thickness=2.0 ** -7) if walls else S._walls_dict(0)
    return sc


# ----------------------------------------------------------------------------------------------
# contact law
# ----------------------------------------------------------------------------------------------
def _lc(vc, n, delta, dt, kn, ks, mu, e, effM, spring=(0, 0, 0)):
    lib = oracle_lib()
    s = D3(*spring)
    F = D3()
    lib.lsdem_oracle_linear_contact(D3(*vc), D3(*n), C.c_double(delta), C.c_double(dt), C.c_double(kn),
                                    C.c_double(ks), C.c_double(mu), C.c_double(e), C.c_double(effM), s, F)
    return np.array(F[:]), np.array(s[:])


def test_normal_force_at_rest():
    F, s = _lc((0, 0, 0), (0, 0, 1), 1e-6, 1e-5, 5e4, 2e4, 0.5, 0.5, 1e-3)
    assert F[2] == pytest.approx(5e4 * 1e-6, rel=1e-15) and F[0] == 0 and F[1] == 0


def test_no_force_without_overlap():
    F, s = _lc((1, 0, 0), (0, 0, 1), 0.0, 1e-5, 5e4, 2e4, 0.5, 0.5, 1e-3, spring=(1e-6, 0, 0))
    assert np.all(F == 0) and np.all(s == 0)


def test_coulomb_cap():
    n = np.array([0.0, 0.0, 1.0])
    F, s = _lc((3.0, 0.5, 0.0), n, 2e-6, 1e-5, 5e4, 2e4, 0.3, 0.7, 1e-3, spring=(1e-5, 0, 0))
    Fn = np.dot(F, n) * n
    Ft = F - Fn
    assert np.linalg.norm(Ft) == pytest.approx(0.3 * np.linalg.norm(Fn), rel=1e-12)


def test_viscous_tangential_without_friction():
    # mu = 0: the spring is zero but the tangential dashpot still acts (reference behaviour)
    F, s = _lc((1.0, 0, 0), (0, 0, 1), 1e-6, 1e-5, 5e4, 2e4, 0.0, 0.5, 1e-3)
    assert np.all(s == 0) and F[0] < 0


@pytest.mark.parametrize("e", [1.0, 0.5, 0.2])
def test_restitution_and_duration_on_wall(S, e):
    """Sphere dropped (no gravity) onto a plane with one node facing it: exit/entry speed = e and
    contact time = pi / (w_n sqrt(1 - b^2)), b = -ln e / sqrt(ln^2 e + pi^2)."""
    r, m, kn = 5e-3, 1.3e-3, 1e5
    sh = sphere_shape(r, 24, r / 10, axis_nodes(r))
    wn = np.sqrt(kn / m)
    b = -np.log(e) / np.sqrt(np.log(e) ** 2 + np.pi ** 2)
    tc = np.pi / (wn * np.sqrt(1 - b * b))
    dt = tc / 2000
    v0 = 0.2
    sc = one_particle_scene(S, sh, [(0, 0, r + 0.5 * v0 * dt)], [(0, 0, -v0)], m, kn, e, dt=dt)
    sim = sc.load(oracle_sim(sc.cfg))
    in_contact = 0
    for _ in range(4000):
        sim.step(1)
        c = sim.contacts(1)
        if len(c["master"]):
            in_contact += 1
        elif in_contact:
            break
    v = sim.get_particles()["vel"][0, 2]
    assert v / v0 == pytest.approx(e, rel=2e-3)
    assert in_contact * dt == pytest.approx(tc, abs=3 * dt)


def test_head_on_restitution(S):
    r, m, kn, e = 5e-3, 1.3e-3, 1e5, 0.6
    sh = sphere_shape(r, 24, r / 10, axis_nodes(r))
    # pair stiffness kn/2, effective mass m/2: same natural frequency as one body on a wall
    wn = np.sqrt(kn / m)
    b = -np.log(e) / np.sqrt(np.log(e) ** 2 + np.pi ** 2)
    tc = np.pi / (wn * np.sqrt(1 - b * b))
    dt = tc / 2000
    v0 = 0.1
    sc = one_particle_scene(S, sh, [(0, 0, 0), (0, 0, 2 * r + v0 * dt)], [(0, 0, v0 / 2), (0, 0, -v0 / 2)], m, kn, e,
                            dt=dt, walls=False)
    sim = sc.load(oracle_sim(sc.cfg))
    started = False
    for _ in range(4000):
        sim.step(1)
        n = len(sim.contacts(0)["master"])
        started |= n > 0
        if started and n == 0:
            break
    v = sim.get_particles()["vel"][:, 2]
    assert (v[1] - v[0]) / v0 == pytest.approx(e, rel=2e-3)
    F = sim.get_particles()["force"]
    assert np.allclose(F.sum(axis=0), 0.0, atol=1e-12)


# ----------------------------------------------------------------------------------------------
# integration
# ----------------------------------------------------------------------------------------------
def test_free_flight_and_rotation(S):
    r, m = 5e-3, 1.3e-3
    sh = sphere_shape(r, 12, r / 4, axis_nodes(r))
    g = (0.3, -1.0, -9.81)
    w0 = np.array([3.0, -2.0, 5.0])
    sc = one_particle_scene(S, sh, [(0, 0, 0.01)], [(0.1, 0.2, 0.5)], m, 1e5, 0.5, g=g, dt=1e-4, walls=False)
    sc.particles["angvel"][:] = w0
    sim = sc.load(oracle_sim(sc.cfg))
    k = 200
    sim.step(k)
    st = sim.get_particles()
    t = k * 1e-4
    x = np.array([0, 0, 0.01]) + np.array([0.1, 0.2, 0.5]) * t + 0.5 * np.array(g) * t * t
    assert np.allclose(st["pos"][0], x, rtol=0, atol=1e-14)
    assert np.allclose(st["angvel"][0], w0, rtol=0, atol=0)
    assert abs(np.linalg.norm(st["quat"][0]) - 1) < 1e-14
    # first-order quaternion steps: rotation angle |w| t up to O(dt)
    q = st["quat"][0]
    ang = 2 * np.arctan2(np.linalg.norm(q[1:]), q[0])
    assert ang == pytest.approx(np.linalg.norm(w0) * t, rel=1e-3)
    assert np.allclose(q[1:] / np.linalg.norm(q[1:]), w0 / np.linalg.norm(w0), atol=1e-9)


def test_fixed_particle_not_integrated(S):
    r = 5e-3
    sh = sphere_shape(r, 12, r / 4, axis_nodes(r))
    sc = one_particle_scene(S, sh, [(0, 0, 0.01)], [(0, 0, 0)], 1.0, 1e5, 0.5, g=(0, 0, -9.81), walls=False)
    sc.particles["inv_mass"][:] = 0.0
    sim = sc.load(oracle_sim(sc.cfg))
    sim.step(10)
    assert np.all(sim.get_particles()["vel"] == 0)


# ----------------------------------------------------------------------------------------------
# interpolation
# ----------------------------------------------------------------------------------------------
def _interp(c, x, y, z, quirk):
    lib = oracle_lib()
    v = C.c_double()
    g = D3()
    lib.lsdem_oracle_interp((C.c_double * 8)(*c), C.c_double(x), C.c_double(y), C.c_double(z), C.c_int(quirk),
                            C.byref(v), g)
    return v.value, np.array(g[:])


def test_linear_field_exact():
    a, b, c, d = 0.25, -1.5, 0.75, 2.0
    corners = [a + b * i + c * j + d * k for k in (0, 1) for j in (0, 1) for i in (0, 1)]
    for (x, y, z) in [(0.1, 0.2, 0.3), (0.9, 0.5, 0.05), (0.0, 0.0, 0.0)]:
        v, g = _interp(corners, x, y, z, 1)
        assert v == pytest.approx(a + b * x + c * y + d * z, abs=1e-15)
        assert np.allclose(g, [b, c, d], atol=1e-15)


def test_int_x_gradient_kat():
    # only phi111 = 1.  Exact gradient at (x,y,z) = (y z, x z, x y); the reference truncates x
    # to 0 inside the gradient, giving (y z, 0, 0).
    corners = [0, 0, 0, 0, 0, 0, 0, 1.0]
    v, g = _interp(corners, 0.7, 0.2, 0.3, 1)
    assert v == pytest.approx(0.7 * 0.2 * 0.3, abs=1e-16)
    assert np.allclose(g, [0.2 * 0.3, 0.0, 0.0], atol=1e-16)
    v, g = _interp(corners, 0.7, 0.2, 0.3, 0)
    assert np.allclose(g, [0.2 * 0.3, 0.7 * 0.3, 0.7 * 0.2], atol=1e-16)


# ----------------------------------------------------------------------------------------------
# mass properties
# ----------------------------------------------------------------------------------------------
def test_sphere_mass_properties():
    """Smoothed-Heaviside mass converges to rho V at second order in the spacing; the
    inertia to (2/5) m r^2; off-diagonals vanish and the inverse is the matrix inverse."""
    lib = oracle_lib()
    import lsdem_amd.lsdem as L
    r, rho = 5e-3, 2500.0
    m0 = 4 / 3 * np.pi * r ** 3 * rho
    errs = []
    for n, h in ((40, 0.5e-3), (80, 0.25e-3)):
        sdf, o, hh, _ = sphere_shape(r, n, h, np.zeros((0, 3)))
        m, c, I, Ii = L.mass_properties(lib, sdf, o, hh, rho, prefix="lsdem_oracle_")
        errs.append(abs(m - m0) / m0)
        assert m == pytest.approx(m0, rel=2e-2)
        assert np.allclose(c, 0, atol=1e-12)
        assert I[0] == pytest.approx(0.4 * m0 * r * r, rel=4e-2)
        assert np.allclose(I[3:], 0, atol=1e-18)
        M = np.array([[I[0], I[3], I[4]], [I[3], I[1], I[5]], [I[4], I[5], I[2]]])
        assert np.allclose(np.linalg.inv(M)[0, 0], Ii[0], rtol=1e-12)
    assert errs[1] < errs[0] / 3.0


# ----------------------------------------------------------------------------------------------
# cell table
# ----------------------------------------------------------------------------------------------
def test_hash_cells_bruteforce():
    lib = oracle_lib()
    rng = np.random.default_rng(1)
    n = 500
    mn, mx = np.array([0.0, -0.1, 0.2]), np.array([0.3, 0.2, 0.35])
    pos = rng.uniform(mn - 0.02, mx + 0.02, size=(n, 3))
    cell = 0.04
    keys = np.zeros(n, np.int32)
    srt = np.zeros(n, np.int32)
    cap = 4096
    st = np.zeros(cap, np.int32)
    en = np.zeros(cap, np.int32)
    nc = C.c_int32()
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    assert lib.lsdem_oracle_hash_cells(D3(*mn), D3(*mx), C.c_double(cell), C.c_int32(n), P(pos), P(keys), P(srt),
                                       C.c_int32(cap), P(st), P(en), C.byref(nc)) == 0
    d = mx - mn
    size = np.where(d > 2 * cell, (d / cell).astype(int), 2)
    inv = size / d
    inside = np.all((pos >= mn) & (pos < mx), axis=1)
    cxyz = ((pos - mn) * inv).astype(int)
    kb = np.where(inside, cxyz[:, 2] * size[1] * size[0] + cxyz[:, 1] * size[0] + cxyz[:, 0], -1)
    assert np.array_equal(kb, keys)
    order = np.argsort(kb, kind="stable")
    assert np.array_equal(order, srt)
    ncell = int(np.prod(size))
    assert nc.value == ncell
    for h in range(ncell):
        idx = np.nonzero(kb[order] == h)[0]
        if len(idx):
            assert st[h] == idx[0] and en[h] == idx[-1] + 1
        else:
            assert st[h] == -1 and en[h] == -1


# ----------------------------------------------------------------------------------------------
# contact sets
# ----------------------------------------------------------------------------------------------
def rotmat(q):
    q0, q1, q2, q3 = q
    return np.array([[1 - 2 * (q2 * q2 + q3 * q3), 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)],
                     [2 * (q1 * q2 + q0 * q3), 1 - 2 * (q1 * q1 + q3 * q3), 2 * (q2 * q3 - q0 * q1)],
                     [2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 1 - 2 * (q1 * q1 + q2 * q2)]])


def brute_pp(sc, st, margin=1e-10):
    """All (node of A, B) with A < B, B inside the domain, |xw - xB| <= rB, xw inside B's grid,
    trilinear phi <= 0, computed with rotation matrices.  Returns (sure set, borderline set)."""
    P = sc.particles
    shapes = sc.shapes
    mn, mx = np.array(sc.cfg["domain_min"]), np.array(sc.cfg["domain_max"])
    pos, quat = st["pos"], st["quat"]
    n = len(pos)
    node0 = np.concatenate([[0], np.cumsum([len(shapes[s]["nodes"]) for s in P["shape"]])])
    R = np.array([rotmat(q) for q in quat])
    rad = np.array([np.linalg.norm(shapes[s]["nodes"] * P["scale"][i], axis=1).max() for i, s in enumerate(P["shape"])])
    inside = np.all((pos >= mn) & (pos < mx), axis=1)
    sure, border = set(), set()
    for A in range(n):
        SA = shapes[P["shape"][A]]
        xw = (R[A] @ (SA["nodes"] * P["scale"][A]).T).T + pos[A]
        for B in range(A + 1, n):
            if not inside[B]:
                continue
            SB = shapes[P["shape"][B]]
            rel = xw - pos[B]
            d2 = (rel * rel).sum(1)
            loc = rel @ R[B]
            s = P["scale"][B]
            g = (loc - SB["origin"] * s) / (SB["spacing"] * s)
            i0 = np.floor(g).astype(int)
            nz, ny, nx = SB["sdf"].shape
            ok = np.all(i0 >= 0, 1) & (i0[:, 0] < nx - 1) & (i0[:, 1] < ny - 1) & (i0[:, 2] < nz - 1)
            f = g - i0
            phi = np.full(len(xw), np.inf)
            for k in np.nonzero(ok)[0]:
                a, b, c = i0[k]
                cube = SB["sdf"][c:c + 2, b:b + 2, a:a + 2].astype(np.float64) * s
                x, y, z = f[k]
                phi[k] = ((cube[0, 0, 0] * (1 - x) + cube[0, 0, 1] * x) * (1 - y) +
                          (cube[0, 1, 0] * (1 - x) + cube[0, 1, 1] * x) * y) * (1 - z) + \
                         ((cube[1, 0, 0] * (1 - x) + cube[1, 0, 1] * x) * (1 - y) +
                          (cube[1, 1, 0] * (1 - x) + cube[1, 1, 1] * x) * y) * z
            r2 = rad[B] ** 2
            for k in range(len(xw)):
                if not ok[k] or d2[k] > r2 * (1 + margin):
                    continue
                key = (node0[A] + k, B)
                if abs(d2[k] - r2) <= margin * r2 or abs(phi[k]) <= margin * s * SB["spacing"]:
                    border.add(key)
                elif phi[k] <= 0:
                    sure.add(key)
    return sure, border


def test_pp_set_bruteforce(S):
    sc = S.scene_c1(n=27, lattice=9.6e-3, jitter=0.4e-3, walls=False)
    sim = sc.load(oracle_sim(sc.cfg))
    st = sim.get_particles()
    sim.step(1)
    st = sim.get_particles()
    c = sim.contacts(0)
    got = set(zip(c["master"].tolist(), c["slave"].tolist()))
    sure, border = brute_pp(sc, st)
    assert len(sure) > 50
    assert sure <= got
    assert got <= (sure | border)


def test_plane_wall_overlap_exact(S):
    # node 1 um below the plane z = 0: overlap 1e-6 and normal +z, exactly representable path
    r = 5e-3
    sh = sphere_shape(r, 12, r / 4, axis_nodes(r))
    sc = one_particle_scene(S, sh, [(0.001, -0.002, r - 1e-6)], [(0, 0, 0)], 1e-3, 1e5, 0.5)
    sim = sc.load(oracle_sim(sc.cfg))
    sim.step(1)
    c = sim.contacts(1)
    assert len(c["master"]) == 1 and c["master"][0] == 5
    assert c["overlap"][0] == pytest.approx(1e-6, rel=1e-9)
    assert np.allclose(c["normal"][0], [0, 0, 1], atol=1e-15)


@pytest.mark.parametrize("periodic", [False, True])
def test_literal_equals_pairs(S, periodic):
    if periodic:
        sc = S.scene_c2(n=150, n_types=2, n_grid=20, n_nodes=96, cell_xy=0.04, gap=-0.12)
    else:
        sc = S.scene_c1(n=40, lattice=9.6e-3, jitter=0.4e-3)
    a = sc.load(oracle_sim(sc.cfg, literal=True))
    b = sc.load(oracle_sim(sc.cfg, literal=False))
    for _ in range(3):
        a.step(1)
        b.step(1)
        for l in (0, 1, 2, 3, 4):
            ca, cb = a.contacts(l), b.contacts(l)
            assert np.array_equal(ca["master"], cb["master"]) and np.array_equal(ca["slave"], cb["slave"])
        pa, pb = a.get_particles(), b.get_particles()
        assert np.allclose(pa["force"], pb["force"], rtol=0, atol=1e-9 * np.abs(pb["force"]).max())


@pytest.mark.parametrize("literal", [False, True])
def test_threads_equal_serial(S, literal):
    """The OpenMP variant of the detection loops (the all-cores CPU baseline) gives the serial
    oracle's lists, springs, forces and counters bit for bit."""
    sc = S.scene_c2(n=150, n_types=2, n_grid=20, n_nodes=96, cell_xy=0.04, gap=-0.12)
    a = sc.load(oracle_sim(sc.cfg, literal=literal))
    b = sc.load(oracle_sim(sc.cfg, literal=literal, threads=4))
    for _ in range(3):
        a.step(2)
        b.step(2)
        for l in (0, 1, 2, 3, 4):
            ca, cb = a.contacts(l), b.contacts(l)
            for k in ca:
                assert np.array_equal(ca[k], cb[k]), (l, k)
        pa, pb = a.get_particles(), b.get_particles()
        for k in ("pos", "vel", "quat", "angvel", "force", "torque"):
            assert np.array_equal(pa[k], pb[k]), k
    assert a.counters() == b.counters()


def test_ghost_third_law(S):
    """Without walls, particle-particle and ghost contacts only exchange internal forces: the
    total force on all particles vanishes (each ghost contact pushes A and the real source j)."""
    sc = S.scene_c2(n=150, n_types=2, n_grid=20, n_nodes=96, cell_xy=0.04, gap=-0.12)
    sc.walls = S._walls_dict(0)
    sc.cfg["gravity"] = (0.0, 0.0, 0.0)
    sim = sc.load(oracle_sim(sc.cfg))
    sim.step(1)
    F = sim.get_particles()["force"]
    assert len(sim.contacts(2)["master"]) + len(sim.contacts(3)["master"]) > 0
    assert np.abs(F.sum(axis=0)).max() <= 1e-10 * np.abs(F).sum()


# ----------------------------------------------------------------------------------------------
# bonds
# ----------------------------------------------------------------------------------------------
def test_bond_axial_stiffness(S):
    r = 5e-3
    sh = sphere_shape(r, 12, r / 4, axis_nodes(r))
    E, rb, L0 = 1e8, 1e-3, 2e-3
    sc = one_particle_scene(S, sh, [(0, 0, 0), (0, 0, 2 * r + L0)], [(0, 0, 0), (0, 0, 0)], 1e-3, 0.0, 1.0,
                            walls=False, dt=1e-7)
    sc.bonds = [(0, 1, (0, 0, r + L0 / 2), (0, 0, -1), rb, L0, E, 0.3)]
    sim = sc.load(oracle_sim(sc.cfg))
    # bond normal points from slave to master (as contact normals do); stretch by 1 um
    st = sim.get_particles()
    st["pos"][1, 2] += 1e-6
    sim.set_particles(pos=st["pos"])
    sim.step(1)
    F = sim.get_particles()["force"]
    k = E * np.pi * rb * rb / L0
    assert F[0, 2] == pytest.approx(k * 1e-6, rel=1e-4)
    assert F[1, 2] == pytest.approx(-k * 1e-6, rel=1e-4)
    b = sim.bonds()
    assert b["active"][0] == 1
    # tensile break: sigma = F / A; set strength just below and above the stress
    for strength, expect in ((0.9 * k * 1e-6 / (np.pi * rb * rb), 0), (1.1 * k * 1e-6 / (np.pi * rb * rb), 1)):
        sc2 = one_particle_scene(S, sh, [(0, 0, 0), (0, 0, 2 * r + L0)], [(0, 0, 0), (0, 0, 0)], 1e-3, 0.0, 1.0,
                                 walls=False, dt=1e-7)
        sc2.bonds = [(0, 1, (0, 0, r + L0 / 2), (0, 0, -1), rb, L0, E, 0.3, strength)]
        s2 = sc2.load(oracle_sim(sc2.cfg))
        st2 = s2.get_particles()
        st2["pos"][1, 2] += 1e-6
        s2.set_particles(pos=st2["pos"])
        s2.step(1)
        assert s2.bonds()["active"][0] == expect
        # the force is still applied on the breaking step
        assert s2.get_particles()["force"][0, 2] > 0


def test_bond_rejects_invalid(S):
    import lsdem_amd.lsdem as L
    r = 5e-3
    sh = sphere_shape(r, 12, r / 4, axis_nodes(r))
    sc = one_particle_scene(S, sh, [(0, 0, 0), (0, 0, 0.02)], [(0, 0, 0), (0, 0, 0)], 1e-3, 0.0, 1.0, walls=False)
    sim = oracle_sim(sc.cfg)
    for s in sc.shapes:
        sim.add_shape(s["sdf"], s["origin"], s["spacing"], s["nodes"])
    P = sc.particles
    sim.add_particles(P["shape"], P["scale"], P["pos"], P["quat"], P["vel"], P["angvel"], P["inv_mass"],
                      P["inv_inertia"], P["kn"], P["ks"], P["mu"], P["e"])
    for bad in [dict(normal=(0, 0, 0)), dict(radius=0.0), dict(length=-1.0), dict(youngs=0.0), dict(poisson=0.5)]:
        args = dict(i=0, j=1, point=(0, 0, 0.01), normal=(0, 0, 1), radius=1e-3, length=1e-3, youngs=1e8, poisson=0.3)
        args.update(bad)
        with pytest.raises(L.LSDEMError):
            sim.add_bond(**args)


# ----------------------------------------------------------------------------------------------
# external-force hook (LSSolver::addExternalForceTorque, LSSolver.h:508,1009)
# ----------------------------------------------------------------------------------------------
def _view_array(ptr, n, k=3):
    return np.ctypeslib.as_array((C.c_double * (k * n)).from_address(ptr)).reshape(n, k)


def test_external_force_hook_time_and_placement(S):
    """The hook sees the summed forces of the step, its additions drive the second half-kick of
    that step (and the first of the next one), and time is the step's start time (sum of dt)."""
    r = 5e-3
    sh = sphere_shape(r, 12, r / 4, axis_nodes(r))
    g = (0.0, 0.0, -9.81)
    sc = one_particle_scene(S, sh, [(0, 0, 0.01)], [(0, 0, 0)], 1.0, 1e5, 0.5, g=g, dt=1e-4, walls=False)
    sim = sc.load(oracle_sim(sc.cfg))
    times = []

    def hook(time, view):
        times.append(time)
        F = _view_array(view.force, view.num)
        assert np.all(F == 0.0)  # nothing else acts on the particle
        F[0] = (0.0, 0.0, 9.81)  # cancels gravity for a unit mass
    sim.set_external_force(hook)
    sim.step(50)
    st = sim.get_particles()
    # the first half-kick of step 1 still sees the initial zero force; every kick after that sees
    # gravity cancelled, so v stays at -g dt/2 and x moves linearly
    vz = 0.0 + (0.0 * 1.0 + -9.81) * (0.5 * 1e-4)
    z = 0.01
    for _ in range(50):
        z += 1e-4 * vz
    assert np.array_equal(st["vel"][0], [0.0, 0.0, vz])
    assert np.array_equal(st["pos"][0], [0.0, 0.0, z])
    t = 0.0
    want = []
    for _ in range(50):
        want.append(t)
        t += 1e-4
    assert times == want
    assert sim.time() == t
    # without the hook the particle falls
    sim.set_external_force(None)
    sim.step(10)
    assert sim.get_particles()["vel"][0][2] < 0


def numpy_damping(view, c):
    """FunDEM/apps/tutorial/globalDamping.cu:4-17 in numpy: F -= c|F| v/|v|, T -= c|T| w/|w|."""
    n = view.num
    for fp, vp in ((view.force, view.velocity), (view.torque, view.angular_velocity)):
        F, v = _view_array(fp, n), _view_array(vp, n)
        l2 = (v * v).sum(1)
        nv = np.where(l2[:, None] > 0, v * (1.0 / np.sqrt(np.where(l2 > 0, l2, 1.0)))[:, None], 0.0)
        F -= c * np.sqrt((F * F).sum(1))[:, None] * nv


def _drop_on_wall(S, hook=None, e=0.8):
    r, m, kn = 5e-3, 1.3e-3, 1e5
    sh = sphere_shape(r, 24, r / 10, axis_nodes(r))
    dt = 1e-6
    v0 = 0.2
    sc = one_particle_scene(S, sh, [(0, 0, r + 0.5 * v0 * dt)], [(0, 0, -v0)], m, kn, e, dt=dt)
    sim = sc.load(oracle_sim(sc.cfg))
    if hook is not None:
        sim.set_external_force(hook)
    seen = 0
    for _ in range(2000):
        sim.step(1)
        if len(sim.contacts(1)["master"]):
            seen += 1
        elif seen:
            break
    return sim.get_particles()["vel"][0, 2] / v0


def test_damping_hook_dissipates(S):
    """The tutorial's global damping (force -= c |F| v/|v|) takes energy out of a wall bounce, and
    the hook sees the contact force of the step (equal to the force reported afterwards)."""
    forces = []

    def record(t, view):
        forces.append(_view_array(view.force, view.num).copy())
    plain = _drop_on_wall(S, hook=record)
    assert any(np.abs(f).max() > 0 for f in forces)
    damped = _drop_on_wall(S, hook=lambda t, v: numpy_damping(v, 0.2))
    assert plain == pytest.approx(0.8, rel=1e-2)
    assert 0 < damped < plain - 0.05
