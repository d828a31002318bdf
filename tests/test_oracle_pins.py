"""Independent pins of the oracle functions that the closed forms of test_oracle.py leave open.

Each test computes its expectation from the physics or from a different formulation (rotation
matrices, numpy clips, beam theory), never by retyping the oracle's expression, and is chosen so
that a plausible mistake fails it:

  function (oracle)                         pinned by                                  a mistake it catches
  makeGhost / ghostKey (PeriodicBoundary.h:215-289, mySpatialGrid.cu:87-152,214-240)
                                            image offsets and keys recomputed with      (+x,-y) image added; image
                                            numpy floor/clip; ghost contact sets by    in the wrong column; z tested
                                            brute force over all (node, image) pairs  against maxBound.z; no clamp
  kick / rotateInvInertia (myMat.h:107-158) one torque kick of a rotated anisotropic   R^T I^-1 R instead of
                                            body vs R diag(I^-1) R^T T h               R I^-1 R^T
  integrateSlidingSpring (contactKernel.cuh:15-21)
                                            a spring with a normal component comes out  missing rescale; missing
                                            in the tangent plane with its old length    projection; sign
  vbond (contactKernel.cuh:138-199), add (VBondedInteraction.h:133-167)
                                            Euler-Bernoulli stiffnesses of a clamped    wrong B2, B3 or B4; swapped
                                            beam: shear 12EI/L^3, end rotation 4EI/L,  frames; sign of the
                                            torsion GJ/L; Mohr-Coulomb threshold with   friction term
                                            compression strengthening the bond
  prime (LSSolver.h:996-998, 1061-1066)     a particle starting outside the periodic    wrap one step late
                                            cell steps exactly as its wrapped copy
"""
import ctypes as C

import numpy as np
import pytest

from oracle_binding import oracle_lib, oracle_sim
from test_oracle import D3, _lc, axis_nodes, one_particle_scene, rotmat, sphere_shape


def _ghosts(sim, s, n):
    lib = oracle_lib()
    pos = np.zeros((n, 3))
    key = np.zeros(n, np.int32)
    rc = lib.lsdem_oracle_ghosts(sim.ctx, C.c_int32(s), C.c_int32(n), pos.ctypes.data_as(C.c_void_p),
                                 key.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return pos, key


def _grid(sim):
    lib = oracle_lib()
    mn, mx, inv = D3(), D3(), D3()
    size = (C.c_int32 * 3)()
    lib.lsdem_oracle_grid(sim.ctx, 0, mn, mx, inv, size)
    return np.array(mn[:]), np.array(mx[:]), np.array(inv[:]), np.array(size[:])


def _sphere_cloud_scene(S, pos, r=4e-3, dom=((0.0, 0.0, 0.0), (0.1, 0.05, 0.1)), quirks=7, rng=None):
    """Spheres (12 nodes) at the given positions in a periodic x/y domain, no walls, no gravity."""
    nodes = np.vstack([axis_nodes(r), r * np.array([[1, 1, 1], [-1, -1, 1], [1, -1, -1], [-1, 1, -1],
                                                    [1, 1, -1], [-1, -1, -1]]) / np.sqrt(3.0)])
    sh = sphere_shape(r, 14, r / 5, nodes)
    sc = S.Scene(cfg={"domain_min": dom[0], "domain_max": dom[1], "gravity": (0, 0, 0), "dt": 1e-6,
                      "periodic_x": 1, "periodic_y": 1, "quirks": quirks})
    sc.shapes.append({"sdf": sh[0], "origin": sh[1], "spacing": sh[2], "nodes": sh[3]})
    n = len(pos)
    P = S._particles(n)
    P["pos"][:] = pos
    P["quat"][:] = S.random_quaternions(rng or np.random.default_rng(3), n)
    P["inv_mass"][:] = 1.0 / 2e-4
    P["inv_inertia"][:] = (1.0 / (0.4 * 2e-4 * r * r),) * 3 + (0, 0, 0)
    P["kn"][:] = 1e5
    P["ks"][:] = 4e4
    P["mu"][:] = 0.5
    P["e"][:] = 0.5
    sc.particles = P
    sc.walls = S._walls_dict(0)
    return sc


def _expected_images(pos, mn, mx, inv, size, quirk_z):
    """Image positions and keys of the three ghost sets, by floor/clip in numpy."""
    L = mx - mn
    cell = np.floor((pos - mn) * inv).astype(int)
    shifts = [np.where(cell[:, 0:1] == 0, [L[0], 0.0, 0.0], 0.0),
              np.where(cell[:, 1:2] == 0, [0.0, L[1], 0.0], 0.0),
              np.where((cell[:, 0:1] == 0) & (cell[:, 1:2] == 0), [L[0], L[1], 0.0], 0.0)]
    out = []
    zmax = mx[1] if quirk_z else mx[2]
    for sh in shifts:
        img = pos + sh
        outside = ((img[:, 0] < mn[0]) | (img[:, 0] >= mx[0]) | (img[:, 1] < mn[1]) | (img[:, 1] >= mx[1]) |
                   (img[:, 2] < mn[2]) | (img[:, 2] >= zmax))
        c = np.floor((img - mn) * inv)
        c = np.clip(c, 0, size - 1).astype(int)
        # a coordinate beyond the top of the (z-vs-maxY) range clamps to the top layer
        c[:, 2] = np.where(img[:, 2] >= zmax, size[2] - 1, c[:, 2])
        key = np.where(outside, c[:, 2] * size[1] * size[0] + c[:, 1] * size[0] + c[:, 0], -1)
        out.append((img, key))
    return out


def _layout(rng, n):
    # particles in every corner column and row, some above maxBound.y in z (the quirk), some inside
    pts = []
    for x in (0.002, 0.015, 0.05, 0.085, 0.098):
        for y in (0.001, 0.012, 0.025, 0.038, 0.049):
            for z in (0.01, 0.03, 0.06, 0.09):
                pts.append((x, y, z))
    pts = np.array(pts)
    return pts[rng.permutation(len(pts))[:n]] + rng.uniform(-5e-4, 5e-4, size=(n, 3))


@pytest.mark.parametrize("quirks", [7, 5])
def test_ghost_images_and_keys_bruteforce(S, quirks):
    rng = np.random.default_rng(11)
    pos = _layout(rng, 90)
    sc = _sphere_cloud_scene(S, pos, quirks=quirks)
    sim = sc.load(oracle_sim(sc.cfg))
    sim.begin_phase()
    st = sim.get_particles()
    mn, mx, inv, size = _grid(sim)
    exp = _expected_images(st["pos"], mn, mx, inv, size, quirk_z=bool(quirks & 2))
    L = mx - mn
    allowed = [np.array(v) for v in ((0, 0, 0), (L[0], 0, 0), (0, L[1], 0), (L[0], L[1], 0))]
    n_corner = 0
    for s in range(3):
        gp, gk = _ghosts(sim, s, len(pos))
        ep, ek = exp[s]
        assert np.array_equal(gp, ep), f"image positions of set {s}"
        assert np.array_equal(gk, ek), f"image keys of set {s}"
        for d in gp - st["pos"]:
            assert any(np.array_equal(d, a) for a in allowed)  # no (+x, -y) image
        n_corner += int(np.sum((gk >= 0) & np.all(gp == st["pos"], axis=1)))
    # the layout exercises shifted images of all three sets and (with the quirk) unshifted ones
    assert all(np.any(_ghosts(sim, s, len(pos))[1] >= 0) for s in range(3))
    if quirks & 2:
        assert n_corner > 0
    else:
        assert n_corner == 0


def _brute_ghost(sc, st, images, grid, margin=1e-10, image_rot=None):
    """(node of A, j) for every node inside image j's grid whose +-1 cell window holds the image's key
    and whose trilinear phi <= 0.  image_rot: the images' rotation matrices (default: the sources').
    Returns (sure, borderline)."""
    mn, mx, inv, size = grid
    P, shapes = sc.particles, sc.shapes
    pos, quat = st["pos"], st["quat"]
    n = len(pos)
    node0 = np.concatenate([[0], np.cumsum([len(shapes[s]["nodes"]) for s in P["shape"]])])
    R = np.array([rotmat(q) for q in quat])
    Rimg = R if image_rot is None else image_rot
    img, key = images
    sure, border = set(), set()
    kc = np.stack([key % size[0], (key // size[0]) % size[1], key // (size[0] * size[1])], axis=1)
    for A in range(n):
        SA = shapes[P["shape"][A]]
        xw = (R[A] @ (SA["nodes"] * P["scale"][A]).T).T + pos[A]
        c = np.floor((xw - mn) * inv).astype(int)
        c = np.clip(c, 0, size - 1)
        for j in range(n):
            if key[j] < 0:
                continue
            win = np.all(np.abs(c - kc[j]) <= 1, axis=1)
            SB = shapes[P["shape"][j]]
            s = P["scale"][j]
            loc = (xw - img[j]) @ Rimg[j]
            g = (loc - SB["origin"] * s) / (SB["spacing"] * s)
            i0 = np.floor(g).astype(int)
            nz, ny, nx = SB["sdf"].shape
            ok = win & np.all(i0 >= 0, 1) & (i0[:, 0] < nx - 1) & (i0[:, 1] < ny - 1) & (i0[:, 2] < nz - 1)
            for k in np.nonzero(ok)[0]:
                a, b, cc = i0[k]
                cube = SB["sdf"][cc:cc + 2, b:b + 2, a:a + 2].astype(np.float64) * s
                x, y, z = g[k] - i0[k]
                phi = ((cube[0, 0, 0] * (1 - x) + cube[0, 0, 1] * x) * (1 - y) +
                       (cube[0, 1, 0] * (1 - x) + cube[0, 1, 1] * x) * y) * (1 - z) + \
                      ((cube[1, 0, 0] * (1 - x) + cube[1, 0, 1] * x) * (1 - y) +
                       (cube[1, 1, 0] * (1 - x) + cube[1, 1, 1] * x) * y) * z
                kk = (node0[A] + k, j)
                if abs(phi) <= margin * s * SB["spacing"]:
                    border.add(kk)
                elif phi <= 0:
                    sure.add(kk)
    return sure, border


def test_ghost_contact_sets_bruteforce(S):
    """Dense periodic packing: every ghost list equals the brute-force set over independently placed
    images (keys, windows, grid bounds and phi by numpy)."""
    rng = np.random.default_rng(5)
    r = 4e-3
    g = np.arange(0.003, 0.1, 7.2e-3)
    gy = np.arange(0.003, 0.05, 7.2e-3)
    X, Y, Z = np.meshgrid(g, gy, [0.01, 0.0172], indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + rng.uniform(-3e-4, 3e-4, size=(X.size, 3))
    sc = _sphere_cloud_scene(S, pos, r=r)
    sim = sc.load(oracle_sim(sc.cfg))
    sim.step(1)
    st = sim.get_particles()
    grid = _grid(sim)
    exp = _expected_images(st["pos"], *grid, quirk_z=True)
    total = 0
    for s in range(3):
        c = sim.contacts(2 + s)
        got = set(zip(c["master"].tolist(), c["slave"].tolist()))
        sure, border = _brute_ghost(sc, st, exp[s], grid)
        assert sure <= got and got <= (sure | border), f"ghost set {s}"
        total += len(sure)
    assert total > 20


def test_torque_kick_rotated_anisotropic(S):
    """One step with a torque applied by the hook: the second half-kick gives
    w = R(q) diag(Iinv) R(q)^T T dt/2 (the first sees the initial zero torque)."""
    r = 5e-3
    sh = sphere_shape(r, 12, r / 4, axis_nodes(r))
    q = np.array([0.8, 0.2, -0.5, 0.26])
    q /= np.linalg.norm(q)
    Iinv = np.array([3.0e6, 1.1e6, 0.4e6])
    dt = 1e-4
    sc = one_particle_scene(S, sh, [(0, 0, 0.02)], [(0, 0, 0)], 1e-3, 0.0, 1.0, walls=False, dt=dt,
                            inv_inertia=np.concatenate([Iinv, [0, 0, 0]]))
    sc.particles["quat"][0] = q
    sim = sc.load(oracle_sim(sc.cfg))
    T = np.array([2e-6, -7e-6, 3e-6])

    def hook(time, view):
        tv = np.ctypeslib.as_array((C.c_double * 3).from_address(view.torque))
        tv[:] = T
    sim.set_external_force(hook)
    sim.step(1)
    st = sim.get_particles()
    R = rotmat(st["quat"][0])
    want = R @ np.diag(Iinv) @ R.T @ T * (0.5 * dt)
    assert np.allclose(st["angvel"][0], want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    # the transposed rotation would give a clearly different vector
    wrong = R.T @ np.diag(Iinv) @ R @ T * (0.5 * dt)
    assert np.abs(wrong - want).max() > 0.1 * np.abs(want).max()


@pytest.mark.parametrize("vs", [(0.0, 0.0, 0.0), (3e-3, -1e-3, 0.0)])
def test_spring_rotated_into_tangent_plane(vs):
    """A spring with a normal component is projected on the tangent plane and rescaled to its old
    length before the tangential motion is added (no Coulomb cap: mu large)."""
    n = np.array([0.0, 0.0, 1.0])
    s0 = np.array([2e-6, 1e-6, 1.5e-6])
    dt = 1e-5
    vs = np.array(vs)
    # vc has only the tangential part vs (zero normal velocity)
    F, s = _lc(tuple(vs), tuple(n), 1e-6, dt, 5e4, 2e4, 1e3, 0.5, 1e-3, spring=tuple(s0))
    s1 = s - vs * dt
    assert abs(np.dot(s1, n)) <= 1e-22
    assert np.linalg.norm(s1) == pytest.approx(np.linalg.norm(s0), rel=1e-14)
    t = s0 - np.dot(s0, n) * n
    assert np.allclose(np.cross(s1, t), 0.0, atol=1e-25) and np.dot(s1, t) > 0


# ----------------------------------------------------------------------------------------------
# bonds: Euler-Bernoulli stiffnesses of the beam between the two bond points
# ----------------------------------------------------------------------------------------------
R_SPH, E_B, NU_B, RB, LB = 5e-3, 1e8, 0.3, 1e-3, 2e-3


def _bond_pair(S, strength=0.0, cohesion=0.0, friction=0.0):
    sh = sphere_shape(R_SPH, 12, R_SPH / 4, axis_nodes(R_SPH))
    # i at the origin, j above it; bond points at z = r and z = r + L (bond normal -z: from j to i)
    sc = one_particle_scene(S, sh, [(0, 0, 0), (0, 0, 2 * R_SPH + LB)], [(0, 0, 0), (0, 0, 0)], 1e-3, 0.0, 1.0,
                            walls=False, dt=1e-9)
    sc.bonds = [(0, 1, (0, 0, R_SPH + LB / 2), (0, 0, -1), RB, LB, E_B, NU_B, strength, cohesion, friction)]
    return sc.load(oracle_sim(sc.cfg))


def _beam():
    A = np.pi * RB ** 2
    I = A * RB ** 2 / 4
    G = E_B / (2 * (1 + NU_B))
    return A, I, 2 * I, G


def _rot_about(axis, theta, point, pos, q):
    """Rotate body (pos, q) by theta about `axis` through `point`."""
    import lsdem_amd.scenes as S
    qr = S.quat_axis_angle(axis, theta)
    Rm = rotmat(qr)
    w0, w1, w2, w3 = qr
    a0, a1, a2, a3 = q
    qn = np.array([w0 * a0 - w1 * a1 - w2 * a2 - w3 * a3, w0 * a1 + w1 * a0 + w2 * a3 - w3 * a2,
                   w0 * a2 - w1 * a3 + w2 * a0 + w3 * a1, w0 * a3 + w1 * a2 - w2 * a1 + w3 * a0])
    return point + Rm @ (pos - point), qn


def _step_moved(sim, pos1, q1):
    st = sim.get_particles()
    st["pos"][1] = pos1
    st["quat"][1] = q1
    sim.set_particles(pos=st["pos"], quat=st["quat"])
    sim.step(1)
    return sim.get_particles()


def test_bond_shear_stiffness(S):
    """j displaced sideways by d (no rotation): |F| = 12 E I / L^3 d, perpendicular to the bond."""
    A, I, J, G = _beam()
    sim = _bond_pair(S)
    d = 1e-7
    st0 = sim.get_particles()
    st = _step_moved(sim, st0["pos"][1] + [d, 0, 0], st0["quat"][1])
    Fi = st["force"][0]
    assert Fi[0] == pytest.approx(12 * E_B * I / LB ** 3 * d, rel=1e-4)  # i pulled towards j's offset
    assert abs(Fi[1]) < 1e-9 * abs(Fi[0])
    assert sum(sim.bonds()["energies"][0]) == pytest.approx(0.5 * 12 * E_B * I / LB ** 3 * d * d, rel=2e-4)


def test_bond_bending_stiffness(S):
    """j rotated by theta about an axis perpendicular to the bond through its bond point: moment
    about that point 4 E I / L theta (clamped beam end), stored energy theta^2 2 E I / L."""
    A, I, J, G = _beam()
    sim = _bond_pair(S)
    th = 1e-4
    st0 = sim.get_particles()
    bp = np.array([0, 0, R_SPH + LB])  # j's bond point
    p1, q1 = _rot_about([0, 1, 0], th, bp, st0["pos"][1], st0["quat"][1])
    st = _step_moved(sim, p1, q1)
    # moment of j's bond end about its bond point: torque - arm x force
    Mj = st["torque"][1] - np.cross(bp - st["pos"][1], st["force"][1])
    assert Mj[1] == pytest.approx(-4 * E_B * I / LB * th, rel=1e-3)
    assert sum(sim.bonds()["energies"][0]) == pytest.approx(0.5 * 4 * E_B * I / LB * th * th, rel=1e-3)


def test_bond_torsion_stiffness(S):
    """j twisted by theta about the bond axis: torque G J / L theta about the axis, no force."""
    A, I, J, G = _beam()
    sim = _bond_pair(S)
    th = 1e-4
    st0 = sim.get_particles()
    p1, q1 = _rot_about([0, 0, 1], th, st0["pos"][1], st0["pos"][1], st0["quat"][1])
    st = _step_moved(sim, p1, q1)
    assert st["torque"][1][2] == pytest.approx(-G * J / LB * th, rel=1e-4)
    assert st["torque"][0][2] == pytest.approx(G * J / LB * th, rel=1e-4)
    assert np.abs(st["force"]).max() < 1e-6 * G * J / LB * th / LB


@pytest.mark.parametrize("compress", [False, True])
def test_bond_mohr_coulomb_break(S, compress):
    """Shear stress tau = 12 E I d / L^3 / A.  Without compression the bond breaks iff tau > c;
    with compression (sigma < 0) the threshold c - mu sigma rises, so a cohesion just below tau no
    longer breaks it."""
    A, I, J, G = _beam()
    d = 2e-7
    tau = 12 * E_B * I / LB ** 3 * d / A
    mu = 0.5
    eps = 1e-8 if compress else 0.0
    sigma = -E_B * A / LB * eps / A  # axial compression eps: sigma = -F/A (compressive, negative)
    for coh, expect in ((0.999 * tau, 1 if compress else 0), (1.001 * tau, 1)):
        sim = _bond_pair(S, cohesion=coh, friction=mu)
        st0 = sim.get_particles()
        _step_moved(sim, st0["pos"][1] + [d, 0, -eps], st0["quat"][1])
        b = sim.bonds()
        assert b["tau_max"][0] == pytest.approx(tau, rel=1e-3)
        if compress:
            assert b["sigma_max"][0] == pytest.approx(sigma, rel=1e-2)
            assert mu * -sigma > 0.002 * tau  # the friction term decides the first case
        assert b["active"][0] == expect


def test_bond_frames_not_renormalised(S):
    """The local frames are stored as rotated (VBondedInteraction.h:161-166): a bond between bodies at
    a generic orientation starts with zero bending and torsion energy to rounding, and Ut of a
    rotated-back frame is exactly the reference expression of the stored vectors."""
    sim = _bond_pair(S)
    sim.step(1)
    e = sim.bonds()["energies"][0]
    assert abs(e[2]) < 1e-12 and abs(e[3]) < 1e-12


# ----------------------------------------------------------------------------------------------
# the pre-loop build: wrap before the first kick
# ----------------------------------------------------------------------------------------------
def test_preloop_wrap(S):
    """A particle that starts one period outside the cell steps exactly as its wrapped copy (the
    pre-loop wrap of LSSolver.h:1063-1066 happens before step 1), positions chosen exactly
    representable so that x - L + L == x."""
    r = 4e-3
    # 0-1 touch inside the cell, 2-3 touch across the x boundary (through a ghost image)
    base = np.array([[0.0625, 0.015625, 0.01], [0.0703125, 0.015625, 0.01], [0.119140625, 0.03125, 0.0121],
                     [0.001953125, 0.03125, 0.0118]])
    moved = base.copy()
    moved[0, 0] -= 0.125  # one period (Lx = 0.125) below the cell
    moved[3, 1] += 0.0625  # one period (Ly = 0.0625) above
    dom = ((0.0, 0.0, 0.0), (0.125, 0.0625, 0.0625))
    out = []
    for p in (base, moved):
        sc = _sphere_cloud_scene(S, p, r=r, dom=dom)
        sc.particles["quat"][:] = (1.0, 0.0, 0.0, 0.0)  # axis nodes point at the neighbours
        sc.cfg["gravity"] = (0.0, 0.0, -9.81)
        sim = sc.load(oracle_sim(sc.cfg))
        sim.step(3)
        out.append((sim.get_particles(), [sim.contacts(k) for k in range(5)]))
    (sa, ca), (sb, cb) = out
    for k in ("pos", "vel", "quat", "angvel", "force"):
        assert np.array_equal(sa[k], sb[k]), k
    for k in range(5):
        assert np.array_equal(ca[k]["master"], cb[k]["master"]) and np.array_equal(ca[k]["slave"], cb[k]["slave"])
    assert len(ca[0]["master"]) > 0 and len(ca[2]["master"]) > 0


def _rz(deg):
    c, s = {90: (0.0, 1.0), 180: (-1.0, 0.0), 270: (0.0, -1.0)}[deg]
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _sector_scene(S, pos, rng=None):
    sc = _sphere_cloud_scene(S, pos, dom=((0.0, 0.0, 0.0), (0.06, 0.06, 0.06)), rng=rng)
    sc.cfg["periodic_x"] = 0
    sc.cfg["periodic_y"] = 0
    return sc


def _sector_sim(sc):
    sim = sc.load(oracle_sim(sc.cfg))
    assert oracle_lib().lsdem_oracle_set_sector(sim.ctx, C.c_int(1)) == 0
    return sim


def test_sector_wrap_images_and_contacts_bruteforce(S):
    """Quarter-sector periodicity (mySpatialGrid.cu:58-85, 154-212): a body that left across x = min
    turns by -90 degrees about the corner's z axis (position, velocity and orientation), one that left
    across y = min by +90; image set 0 is the +90 turn of bodies in the first y cell, set 1 the 180 turn
    of those in the first x and y cells, set 2 the -90 turn of those in the first x cell, with keys only
    outside the box.  Checked with exact quarter-turn matrices in numpy, and the image contact sets by
    brute force over all (node, image) pairs with the images' own orientations."""
    rng = np.random.default_rng(5)
    pts = []
    for x in (0.003, 0.006, 0.03, 0.055):
        for y in (0.002, 0.006, 0.03, 0.056):
            pts.append((x, y, 0.03 + 0.004 * len(pts) % 0.02))
    pos = np.array(pts) + rng.uniform(-4e-4, 4e-4, (len(pts), 3))
    pos[0] = (-0.0015, 0.021, 0.02)   # left across x = min: turned by -90 degrees
    pos[1] = (0.033, -0.0012, 0.04)   # left across y = min: turned by +90
    sc = _sector_scene(S, pos, rng)
    sc.particles["vel"][:2] = ((0.3, -0.2, 0.1), (-0.1, 0.25, 0.0))
    q0 = sc.particles["quat"][:2].copy()
    sim = _sector_sim(sc)
    sim.begin_phase()
    st = sim.get_particles()
    mn, mx, inv, size = _grid(sim)
    for i, deg in ((0, 270), (1, 90)):
        Rz = _rz(deg)
        assert np.allclose(st["pos"][i], mn + Rz @ (pos[i] - mn), rtol=0, atol=1e-15)
        assert np.allclose(st["vel"][i], Rz @ sc.particles["vel"][i], rtol=0, atol=1e-15)
        assert np.allclose(rotmat(st["quat"][i]), Rz @ rotmat(q0[i]), rtol=0, atol=1e-14)
    p = st["pos"]
    cell = np.floor((p - mn) * inv).astype(int)
    sets = [(90, cell[:, 1] == 0), (180, (cell[:, 0] == 0) & (cell[:, 1] == 0)), (270, cell[:, 0] == 0)]
    zmax = mx[1]
    total = 0
    for s, (deg, on) in enumerate(sets):
        Rz = _rz(deg)
        exp = np.where(on[:, None], mn + (p - mn) @ Rz.T, p)
        gp, gk = _ghosts(sim, s, len(p))
        assert np.allclose(gp, exp, rtol=0, atol=1e-15), f"image positions of set {s}"
        outside = ((exp[:, 0] < mn[0]) | (exp[:, 0] >= mx[0]) | (exp[:, 1] < mn[1]) | (exp[:, 1] >= mx[1]) |
                   (exp[:, 2] < mn[2]) | (exp[:, 2] >= zmax))
        c = np.clip(np.floor((exp - mn) * inv), 0, size - 1).astype(int)
        key = np.where(outside, c[:, 2] * size[1] * size[0] + c[:, 1] * size[0] + c[:, 0], -1)
        assert np.array_equal(gk, key), f"image keys of set {s}"
        assert np.any(gk >= 0)
        rot = np.array([(Rz if on[j] else np.eye(3)) @ rotmat(st["quat"][j]) for j in range(len(p))])
        sure, border = _brute_ghost(sc, st, (gp, gk), (mn, mx, inv, size), image_rot=rot)
        got = sim.contacts(2 + s)
        pairs = set(zip(got["master"].tolist(), got["slave"].tolist()))
        assert sure <= pairs and pairs <= sure | border, f"set {s}"
        total += len(pairs)
    assert total > 0
