"""Synthetic scenes of the benchmark configurations (numpy; host-side setup only).

Shapes follow the reference's level-set generators: a superellipsoid SDF grid holds
value/|grad value| of the implicit function (FunDEM/kernel/helper/LevelSetObjectGenerator.h:
122-224, 989-1361), boundary nodes are Fibonacci-sphere directions projected radially onto the
surface (the radial projection of :1011-1044), and mass properties come from a smoothed
Heaviside over the grid (FunDEM/kernel/LSSolver.h:51-163).  Plane walls are thin slab grids
holding phi = z, which trilinear interpolation reproduces exactly.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GOLDEN_ANGLE = np.pi * (3.0 - np.sqrt(5.0))


# ----------------------------------------------------------------------------------------------
# Shapes
# ----------------------------------------------------------------------------------------------
def superellipsoid_value_grad(p: np.ndarray, rx, ry, rz, ee, en):
    """Implicit function f = (|x/rx|^(2/ee) + |y/ry|^(2/ee))^(ee/en) + |z/rz|^(2/en) - 1 and its gradient."""
    ax, ay, az = np.abs(p[..., 0] / rx), np.abs(p[..., 1] / ry), np.abs(p[..., 2] / rz)
    exy, ez, outer = 2.0 / ee, 2.0 / en, ee / en
    xt, yt = ax ** exy, ay ** exy
    s = xt + yt
    f = s ** outer + az ** ez - 1.0
    eps = 1e-30
    so = np.maximum(s, eps) ** (outer - 1.0)
    sgn = lambda v: np.where(v >= 0.0, 1.0, -1.0)
    dsx = exy * np.maximum(ax, eps) ** (exy - 1.0) * sgn(p[..., 0]) / rx
    dsy = exy * np.maximum(ay, eps) ** (exy - 1.0) * sgn(p[..., 1]) / ry
    gx = outer * so * dsx
    gy = outer * so * dsy
    gz = ez * np.maximum(az, eps) ** (ez - 1.0) * sgn(p[..., 2]) / rz
    return f, np.stack([gx, gy, gz], axis=-1)


def superellipsoid_grid(rx, ry, rz, ee, en, n: int, spacing: float):
    """Centred n^3 grid of value/|grad| (float32, indexed [z, y, x]) and its origin."""
    o = -0.5 * (n - 1) * spacing
    ax = o + spacing * np.arange(n)
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    f, g = superellipsoid_value_grad(np.stack([x, y, z], axis=-1), rx, ry, rz, ee, en)
    gn = np.linalg.norm(g, axis=-1)
    sdf = np.where(gn > 1e-12, f / np.where(gn > 1e-12, gn, 1.0), np.where(f <= 0.0, -0.5, 0.5) * spacing)
    return sdf.astype(np.float32), np.array([o, o, o])


def fibonacci_directions(m: int) -> np.ndarray:
    i = np.arange(m)
    z = 1.0 - (2.0 * i + 1.0) / m
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = i * GOLDEN_ANGLE
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def superellipsoid_nodes(rx, ry, rz, ee, en, m: int) -> np.ndarray:
    d = fibonacci_directions(m)
    exy, ez, outer = 2.0 / ee, 2.0 / en, ee / en
    s = (np.abs(d[:, 0] / rx) ** exy + np.abs(d[:, 1] / ry) ** exy) ** outer + np.abs(d[:, 2] / rz) ** ez
    return d * (s ** (-0.5 * en))[:, None]


def plane_wall_shape(half_extent: float, thickness: float):
    """Slab grid around the local plane z = 0 holding phi = z (normal +z).  A power-of-two
    thickness keeps the corner values +-h/2 exact in float32."""
    h = thickness
    n = 2 * int(np.ceil(half_extent / h)) + 1
    o = np.array([-0.5 * (n - 1) * h, -0.5 * (n - 1) * h, -0.5 * h])
    zs = o[2] + h * np.arange(2)
    sdf = np.broadcast_to(zs[:, None, None], (2, n, n)).astype(np.float32).copy()
    return sdf, o, h


def mass_properties(sdf: np.ndarray, origin, h: float, density: float):
    """Mass, centroid, inertia (xx,yy,zz,xy,xz,yz) and inverse inertia, smoothed Heaviside eps = 1.5 cells."""
    phi = sdf.astype(np.float64) / h
    eps = 1.5
    x = -phi / eps
    H = np.where(phi < -eps, 1.0, np.where(phi > eps, 0.0, 0.5 * (1.0 + x + np.sin(np.pi * x) / np.pi)))
    nz, ny, nx = sdf.shape
    zc = origin[2] + h * np.arange(nz)
    yc = origin[1] + h * np.arange(ny)
    xc = origin[0] + h * np.arange(nx)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    mg = density * h ** 3
    m = H.sum() * mg
    c = np.array([(H * X).sum(), (H * Y).sum(), (H * Z).sum()]) * mg / m
    rx, ry, rz = X - c[0], Y - c[1], Z - c[2]
    I = np.array([(H * (ry * ry + rz * rz)).sum(), (H * (rx * rx + rz * rz)).sum(), (H * (rx * rx + ry * ry)).sum(),
                  -(H * rx * ry).sum(), -(H * rx * rz).sum(), -(H * ry * rz).sum()]) * mg
    return m, c, I, inv_sym(I)


def inv_sym(I):
    M = np.array([[I[0], I[3], I[4]], [I[3], I[1], I[5]], [I[4], I[5], I[2]]])
    Mi = np.linalg.inv(M)
    return np.array([Mi[0, 0], Mi[1, 1], Mi[2, 2], Mi[0, 1], Mi[0, 2], Mi[1, 2]])


def random_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random unit quaternions (q0 scalar first), Shoemake's method."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    s1, s2 = np.sqrt(1.0 - u1), np.sqrt(u1)
    a, b = 2 * np.pi * u2, 2 * np.pi * u3
    return np.stack([s2 * np.cos(b), s1 * np.sin(a), s1 * np.cos(a), s2 * np.sin(b)], axis=1)


def quat_axis_angle(axis, angle):
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    return np.concatenate([[np.cos(angle / 2)], np.sin(angle / 2) * axis])


# ----------------------------------------------------------------------------------------------
# Scene container
# ----------------------------------------------------------------------------------------------
@dataclass
class Scene:
    cfg: dict
    shapes: list = field(default_factory=list)      # dicts: sdf, origin, spacing, nodes (+ mass props)
    particles: dict = field(default_factory=dict)
    walls: dict = field(default_factory=dict)
    bonds: list = field(default_factory=list)
    meta: dict = field(default_factory=dict)

    @property
    def n_particles(self):
        return len(self.particles.get("shape", []))

    def load(self, sim):
        for s in self.shapes:
            sim.add_shape(s["sdf"], s["origin"], s["spacing"], s["nodes"])
        p = self.particles
        if self.n_particles:
            sim.add_particles(p["shape"], p["scale"], p["pos"], p["quat"], p["vel"], p["angvel"], p["inv_mass"],
                              p["inv_inertia"], p["kn"], p["ks"], p["mu"], p["e"])
        w = self.walls
        if w and len(w["shape"]):
            sim.add_walls(w["shape"], w["radius"], w["pos"], w["quat"], w["vel"], w["angvel"], w["mu"], w["e"])
        for b in self.bonds:
            sim.add_bond(*b)
        sim.initialize()
        return sim


def _add_shape(shapes, sdf, origin, spacing, nodes, density=None):
    s = {"sdf": sdf, "origin": np.asarray(origin, float), "spacing": float(spacing), "nodes": np.asarray(nodes, float)}
    if density is not None:
        m, c, I, Ii = mass_properties(sdf, origin, spacing, density)
        # shift the local frame to the centroid, as addLSParticle does
        s["origin"] = s["origin"] - c
        s["nodes"] = s["nodes"] - c
        s["mass"], s["inertia"], s["inv_inertia"] = m, I, Ii
        s["density"] = float(density)
    shapes.append(s)
    return len(shapes) - 1


def _walls_dict(n):
    return {"shape": np.zeros(n, np.int32), "radius": np.zeros(n), "pos": np.zeros((n, 3)), "quat": np.zeros((n, 4)),
            "vel": np.zeros((n, 3)), "angvel": np.zeros((n, 3)), "mu": np.zeros(n), "e": np.ones(n)}


def box_walls(shapes, lo, hi, mu, e, thickness=2.0 ** -7):
    """Six inward-facing plane walls on the faces of the box [lo, hi]."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    half = 0.5 * np.max(hi - lo) + thickness
    sdf, o, h = plane_wall_shape(half, thickness)
    sid = _add_shape(shapes, sdf, o, h, np.zeros((0, 3)))
    c = 0.5 * (lo + hi)
    faces = [
        ((c[0], c[1], lo[2]), quat_axis_angle([1, 0, 0], 0.0)),
        ((c[0], c[1], hi[2]), quat_axis_angle([1, 0, 0], np.pi)),
        ((lo[0], c[1], c[2]), quat_axis_angle([0, 1, 0], np.pi / 2)),
        ((hi[0], c[1], c[2]), quat_axis_angle([0, 1, 0], -np.pi / 2)),
        ((c[0], lo[1], c[2]), quat_axis_angle([1, 0, 0], -np.pi / 2)),
        ((c[0], hi[1], c[2]), quat_axis_angle([1, 0, 0], np.pi / 2)),
    ]
    w = _walls_dict(6)
    for k, (p, q) in enumerate(faces):
        w["shape"][k] = sid
        w["radius"][k] = np.sqrt(2.0) * half
        w["pos"][k] = p
        w["quat"][k] = q
        w["mu"][k] = mu
        w["e"][k] = e
    return w


def floor_wall(shapes, center_xy, z, half_extent, mu, e, thickness=2.0 ** -6):
    sdf, o, h = plane_wall_shape(half_extent, thickness)
    sid = _add_shape(shapes, sdf, o, h, np.zeros((0, 3)))
    w = _walls_dict(1)
    w["shape"][0] = sid
    w["radius"][0] = np.sqrt(2.0) * half_extent
    w["pos"][0] = (center_xy[0], center_xy[1], z)
    w["quat"][0] = (1.0, 0.0, 0.0, 0.0)
    w["mu"][0] = mu
    w["e"][0] = e
    return w


def _particles(n):
    return {"shape": np.zeros(n, np.int32), "scale": np.ones(n), "pos": np.zeros((n, 3)), "quat": np.zeros((n, 4)),
            "vel": np.zeros((n, 3)), "angvel": np.zeros((n, 3)), "inv_mass": np.zeros(n),
            "inv_inertia": np.zeros((n, 6)), "kn": np.zeros(n), "ks": np.zeros(n), "mu": np.zeros(n),
            "e": np.ones(n)}


# ----------------------------------------------------------------------------------------------
# Config 1: 1,000 superellipsoids (5/4/3 mm, exponent 4) settling in a 0.1 x 0.1 x 0.3 m box
# ----------------------------------------------------------------------------------------------
def scene_c1(n: int = 1000, seed: int = 20261203, n_grid: int = 40, spacing: float = 5e-4, n_nodes: int = 500,
             lattice: float = 12.5e-3, jitter: float = 0.5e-3, box=(0.1, 0.1, 0.3), dt: float = 1e-5,
             kn: float = 1e5, restitution: float = 0.5, mu: float = 0.5, density: float = 2500.0,
             gravity: float = 9.81, quirks: int = 7, walls: bool = True) -> Scene:
    rng = np.random.default_rng(seed)
    a, b, c, p = 5e-3, 4e-3, 3e-3, 4.0
    ee = en = 2.0 / p
    shapes: list = []
    sdf, org = superellipsoid_grid(a, b, c, ee, en, n_grid, spacing)
    nodes = superellipsoid_nodes(a, b, c, ee, en, n_nodes)
    sid = _add_shape(shapes, sdf, org, spacing, nodes, density)
    S = shapes[sid]
    box = np.asarray(box, float)
    per_row = int(np.floor((box[0] - lattice) / lattice)) + 1
    per_col = int(np.floor((box[1] - lattice) / lattice)) + 1
    P = _particles(n)
    k = np.arange(n)
    ix, iy, iz = k % per_row, (k // per_row) % per_col, k // (per_row * per_col)
    base = np.zeros((n, 3))
    base[:, 0] = lattice * (ix + 0.5) + 0.5 * (box[0] - per_row * lattice)
    base[:, 1] = lattice * (iy + 0.5) + 0.5 * (box[1] - per_col * lattice)
    base[:, 2] = lattice * (iz + 0.5) + 1e-3
    P["pos"] = base + rng.uniform(-jitter, jitter, size=(n, 3))
    P["quat"] = random_quaternions(rng, n)
    P["inv_mass"][:] = 1.0 / S["mass"]
    P["inv_inertia"][:] = S["inv_inertia"]
    P["kn"][:] = kn
    P["ks"][:] = 0.4 * kn
    P["mu"][:] = mu
    P["e"][:] = restitution
    margin = 0.01
    cfg = {"domain_min": (-margin, -margin, -margin), "domain_max": tuple(box + margin),
           "gravity": (0.0, 0.0, -gravity), "dt": dt, "quirks": quirks}
    sc = Scene(cfg=cfg, shapes=shapes, particles=P, meta={"name": "C1", "seed": seed})
    sc.walls = box_walls(shapes, (0, 0, 0), box, mu, restitution) if walls else _walls_dict(0)
    return sc


# ----------------------------------------------------------------------------------------------
# Config 5 building block: one rotating drum (the reference's CylinderWall with the level set
# positive inside, LevelSetObjectGenerator.h:1877-2136, here as an analytic grid)
# ----------------------------------------------------------------------------------------------
def drum_wall_shape(radius: float, depth: float, spacing: float, margin_cells: int = 3):
    """Closed cylinder with axis y, centred at the origin, as a wall grid (float32, [z, y, x]):
    inside the cavity the distance to the nearest face (mantle or end cap), outside minus the
    distance to the cylinder."""
    h = float(spacing)
    nr = 2 * (int(np.ceil(radius / h)) + margin_cells) + 1
    ny = 2 * (int(np.ceil(0.5 * depth / h)) + margin_cells) + 1
    o = np.array([-0.5 * (nr - 1) * h, -0.5 * (ny - 1) * h, -0.5 * (nr - 1) * h])
    z, y, x = np.meshgrid(o[2] + h * np.arange(nr), o[1] + h * np.arange(ny), o[0] + h * np.arange(nr), indexing="ij")
    r, a = np.sqrt(x * x + z * z), np.abs(y)
    inside = (r <= radius) & (a <= 0.5 * depth)
    d_in = np.minimum(radius - r, 0.5 * depth - a)
    d_out = -np.sqrt(np.maximum(r - radius, 0.0) ** 2 + np.maximum(a - 0.5 * depth, 0.0) ** 2)
    return np.where(inside, d_in, d_out).astype(np.float32), o, h


def scene_drum(n: int = 4000, seed: int = 20261206, radius: float = 0.1, depth: float = 0.05, rpm: float = 10.0,
               mu: float = 0.5, restitution: float = 0.5, n_types: int = 4, n_grid: int = 32, n_nodes: int = 300,
               semi_axes=(3e-3, 5e-3), exponents=(2.0, 6.0), wall_spacing: float = 1e-3, dt: float = 1e-5,
               kn: float = 1e5, density: float = 2500.0, gravity: float = 9.81, fill: float = 0.5) -> Scene:
    """A rotating drum (axis y, speed rpm) holding superellipsoids (exponent and semi-axes uniform in
    the given ranges, SDF n_grid^3, n_nodes nodes) on a jittered lattice in its lower part, at most
    `fill` of the cavity's lattice sites.  n is capped by what fits."""
    rng = np.random.default_rng(seed)
    shapes: list = []
    types = []
    for _ in range(n_types):
        ax = rng.uniform(semi_axes[0], semi_axes[1], 3)
        p1, p2 = rng.uniform(exponents[0], exponents[1], 2)
        ee, en = 2.0 / p1, 2.0 / p2
        h = 2.0 * ax.max() / (n_grid - 5)
        sdf, org = superellipsoid_grid(*ax, ee, en, n_grid, h)
        sid = _add_shape(shapes, sdf, org, h, superellipsoid_nodes(*ax, ee, en, n_nodes), density)
        types.append((sid, float(np.linalg.norm(shapes[sid]["nodes"], axis=1).max())))
    rb = max(t[1] for t in types)  # largest bounding radius
    pitch = 2.0 * rb * 1.01
    g = np.arange(-radius + 1.01 * rb, radius - rb, pitch)
    gy = np.arange(-0.5 * depth + 1.01 * rb, 0.5 * depth - rb, pitch)
    X, Y, Z = np.meshgrid(g, gy, g, indexing="ij")
    site = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    site = site[np.sqrt(site[:, 0] ** 2 + site[:, 2] ** 2) <= radius - 1.01 * rb]
    site = site[np.argsort(site[:, 2], kind="stable")]  # lowest first
    n = int(min(n, int(fill * len(site))))
    P = _particles(n)
    kind = rng.integers(0, n_types, n)
    P["pos"] = site[:n] + rng.uniform(-0.002 * pitch, 0.002 * pitch, (n, 3))
    P["quat"] = random_quaternions(rng, n)
    for t, (sid, _) in enumerate(types):
        m = kind == t
        S = shapes[sid]
        P["shape"][m] = sid
        P["inv_mass"][m] = 1.0 / S["mass"]
        P["inv_inertia"][m] = S["inv_inertia"]
    P["kn"][:] = kn
    P["ks"][:] = 0.4 * kn
    P["mu"][:] = mu
    P["e"][:] = restitution
    sdf, o, h = drum_wall_shape(radius, depth, wall_spacing)
    wid = _add_shape(shapes, sdf, o, h, np.zeros((0, 3)))
    w = _walls_dict(1)
    w["shape"][0] = wid
    w["radius"][0] = np.sqrt(radius ** 2 + (0.5 * depth) ** 2)
    w["quat"][0] = (1.0, 0.0, 0.0, 0.0)
    w["angvel"][0] = (0.0, rpm * 2.0 * np.pi / 60.0, 0.0)
    w["mu"][0] = mu
    w["e"][0] = restitution
    m = radius + 4 * wall_spacing
    cfg = {"domain_min": (-m, -0.5 * depth - 4 * wall_spacing, -m), "domain_max": (m, 0.5 * depth + 4 * wall_spacing, m),
           "gravity": (0.0, 0.0, -gravity), "dt": dt}
    return Scene(cfg=cfg, shapes=shapes, particles=P, walls=w,
                 meta={"name": "drum", "seed": seed, "rpm": rpm, "mu": mu, "e": restitution})


def scene_drum_ensemble(counts, rpms, mus, restitutions, seed: int = 20261208, radius: float = 0.1,
                        depth: float = 0.05, n_types: int = 8, n_grid: int = 32, n_nodes: int = 300,
                        semi_axes=(3e-3, 5e-3), exponents=(2.0, 6.0), wall_spacing: float = 1e-3, dt: float = 1e-5,
                        kn: float = 1e5, density: float = 2500.0, gravity: float = 9.81, fill: float = 0.5,
                        gap: float = 0.06, max_count=None) -> Scene:
    """Independent rotating drums in one simulation: drum k (speed rpms[k], friction mus[k],
    restitution restitutions[k]) holds up to counts[k] superellipsoids on a jittered lattice in its
    lower part, as scene_drum does (capped by what `fill` of its lattice sites holds; with max_count,
    the counts are first scaled so that max_count maps to that cap, keeping their spread).  The drums share
    a pool of n_types shapes and one drum-wall shape, and sit on a square array in the x-z plane, their
    grid boxes `gap` apart, so that no particle of one drum ever lies in another drum's grid box: every
    drum evolves exactly as it would alone, and one step of the simulation steps all of them.
    meta["drum_of"] gives each particle's drum, meta["centers"] the drum centres."""
    rng = np.random.default_rng(seed)
    shapes: list = []
    types = []
    for _ in range(n_types):
        ax = rng.uniform(semi_axes[0], semi_axes[1], 3)
        p1, p2 = rng.uniform(exponents[0], exponents[1], 2)
        ee, en = 2.0 / p1, 2.0 / p2
        h = 2.0 * ax.max() / (n_grid - 5)
        sdf, org = superellipsoid_grid(*ax, ee, en, n_grid, h)
        sid = _add_shape(shapes, sdf, org, h, superellipsoid_nodes(*ax, ee, en, n_nodes), density)
        types.append((sid, float(np.linalg.norm(shapes[sid]["nodes"], axis=1).max())))
    rb = max(t[1] for t in types)
    pitch = 2.0 * rb * 1.01
    g = np.arange(-radius + 1.01 * rb, radius - rb, pitch)
    gy = np.arange(-0.5 * depth + 1.01 * rb, 0.5 * depth - rb, pitch)
    X, Y, Z = np.meshgrid(g, gy, g, indexing="ij")
    site = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    site = site[np.sqrt(site[:, 0] ** 2 + site[:, 2] ** 2) <= radius - 1.01 * rb]
    site = site[np.argsort(site[:, 2], kind="stable")]  # lowest first
    wsdf, wo, wh = drum_wall_shape(radius, depth, wall_spacing)
    wid = _add_shape(shapes, wsdf, wo, wh, np.zeros((0, 3)))
    half_box = 0.5 * (wsdf.shape[2] - 1) * wh  # the wall grid's half-width in x and z
    K = len(counts)
    side = int(np.ceil(np.sqrt(K)))
    step = 2.0 * half_box + gap
    centers = np.array([((k % side) * step, 0.0, (k // side) * step) for k in range(K)])
    parts, walls, drum_of = [], _walls_dict(K), []
    cap = int(fill * len(site))
    scale = 1.0 if max_count is None else cap / float(max_count)
    for k in range(K):
        n = int(min(max(1, round(counts[k] * scale)), cap))
        P = _particles(n)
        kind = rng.integers(0, n_types, n)
        P["pos"] = centers[k] + site[:n] + rng.uniform(-0.002 * pitch, 0.002 * pitch, (n, 3))
        P["quat"] = random_quaternions(rng, n)
        for t, (sid, _) in enumerate(types):
            m = kind == t
            S = shapes[sid]
            P["shape"][m] = sid
            P["inv_mass"][m] = 1.0 / S["mass"]
            P["inv_inertia"][m] = S["inv_inertia"]
        P["kn"][:] = kn
        P["ks"][:] = 0.4 * kn
        P["mu"][:] = mus[k]
        P["e"][:] = restitutions[k]
        parts.append(P)
        drum_of.append(np.full(n, k, np.int32))
        walls["shape"][k] = wid
        walls["radius"][k] = np.sqrt(radius ** 2 + (0.5 * depth) ** 2)
        walls["pos"][k] = centers[k]
        walls["quat"][k] = (1.0, 0.0, 0.0, 0.0)
        walls["angvel"][k] = (0.0, rpms[k] * 2.0 * np.pi / 60.0, 0.0)
        walls["mu"][k] = mus[k]
        walls["e"][k] = restitutions[k]
    P = {key: np.concatenate([q[key] for q in parts]) for key in parts[0]}
    m = half_box + wall_spacing
    lo = centers.min(axis=0) - m
    hi = centers.max(axis=0) + m
    lo[1], hi[1] = -0.5 * depth - 4 * wall_spacing, 0.5 * depth + 4 * wall_spacing
    cfg = {"domain_min": tuple(lo), "domain_max": tuple(hi), "gravity": (0.0, 0.0, -gravity), "dt": dt}
    return Scene(cfg=cfg, shapes=shapes, particles=P, walls=walls,
                 meta={"name": "drum ensemble", "seed": seed, "drum_of": np.concatenate(drum_of),
                       "centers": centers, "rpm": list(rpms), "mu": list(mus), "e": list(restitutions)})


def scene_drum_single(ens: Scene, k: int) -> Scene:
    """Drum k of an ensemble as a simulation of its own (same shapes, particles, wall and time step;
    the domain shrunk around it)."""
    sel = ens.meta["drum_of"] == k
    P = {key: v[sel].copy() for key, v in ens.particles.items()}
    W = {key: v[k:k + 1].copy() for key, v in ens.walls.items()}
    c = ens.meta["centers"][k]
    lo, hi = np.array(ens.cfg["domain_min"]), np.array(ens.cfg["domain_max"])
    half = 0.5 * (hi[0] - lo[0] - (ens.meta["centers"][:, 0].max() - ens.meta["centers"][:, 0].min()))
    dlo = np.array([c[0] - half, lo[1], c[2] - half])
    dhi = np.array([c[0] + half, hi[1], c[2] + half])
    cfg = dict(ens.cfg)
    cfg["domain_min"], cfg["domain_max"] = tuple(dlo), tuple(dhi)
    return Scene(cfg=cfg, shapes=ens.shapes, particles=P, walls=W, meta={"name": f"drum {k}"})


# ----------------------------------------------------------------------------------------------
# Config 2: 200,000 superquadrics of 10 types, periodic x/y (0.3 m cell), floor wall
# ----------------------------------------------------------------------------------------------
def unit_superquadric_types(rng, n_types: int, n_grid: int, n_nodes: int):
    """Shapes of unit volume-equivalent radius: exponents U[2,8], axis ratios (1, U[.5,1], U[.5,1])."""
    out = []
    for _ in range(n_types):
        p1, p2 = rng.uniform(2.0, 8.0, 2)
        ee, en = 2.0 / p1, 2.0 / p2
        ax = np.array([1.0, rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0)])
        # normalise to the volume of the unit sphere, volume from a fine point count
        g = np.linspace(-1, 1, 121)
        Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
        f, _ = superellipsoid_value_grad(np.stack([X * ax[0], Y * ax[1], Z * ax[2]], -1), *ax, ee, en)
        vol = (f <= 0).mean() * 8.0 * np.prod(ax)
        ax = ax * (4.0 / 3.0 * np.pi / vol) ** (1.0 / 3.0)
        h = 2.0 * ax.max() / (n_grid - 5)
        sdf, org = superellipsoid_grid(*ax, ee, en, n_grid, h)
        nodes = superellipsoid_nodes(*ax, ee, en, n_nodes)
        out.append((sdf, org, h, nodes, ax, (ee, en)))
    return out


def scene_c2(n: int = 200_000, seed: int = 20261204, n_types: int = 10, n_grid: int = 50, n_nodes: int = 512,
             cell_xy: float = 0.3, median_r: float = 3e-3, sigma: float = 0.25, dt: float = 5e-6, kn: float = 1e5,
             restitution: float = 0.5, mu: float = 0.5, density: float = 2500.0, gravity: float = 9.81,
             gap: float = 0.02, quirks: int = 7) -> Scene:
    """Random-lattice start: particles sorted by size fill square layers whose pitch is the
    largest bounding diameter of the layer (plus `gap`, relative)."""
    rng = np.random.default_rng(seed)
    shapes: list = []
    types = unit_superquadric_types(rng, n_types, n_grid, n_nodes)
    rb = []
    for sdf, org, h, nodes, ax, _ in types:
        sid = _add_shape(shapes, sdf, org, h, nodes, density)
        rb.append(np.linalg.norm(shapes[sid]["nodes"], axis=1).max())
    rb = np.asarray(rb)
    P = _particles(n)
    P["shape"] = rng.integers(0, n_types, n).astype(np.int32)
    P["scale"] = np.exp(np.log(median_r) + sigma * rng.standard_normal(n))
    rbound = rb[P["shape"]] * P["scale"]
    order = np.argsort(-rbound, kind="stable")
    z = 0.0
    k = 0
    while k < n:
        pitch = 2.0 * rbound[order[k]] * (1.0 + gap)
        m = max(1, int(np.floor(cell_xy / pitch)))
        take = order[k:k + m * m]
        idx = np.arange(len(take))
        slack = cell_xy / m - pitch
        P["pos"][take, 0] = (idx % m + 0.5) * (cell_xy / m) + rng.uniform(-0.5, 0.5, len(take)) * slack
        P["pos"][take, 1] = (idx // m + 0.5) * (cell_xy / m) + rng.uniform(-0.5, 0.5, len(take)) * slack
        # the first layer clears the floor by its bounding radius
        zc = z + (0.5 * pitch if z > 0.0 else rbound[order[k]] * (1.0 + abs(gap)))
        P["pos"][take, 2] = zc
        z = zc + 0.5 * pitch
        k += len(take)
    top = z
    P["quat"] = random_quaternions(rng, n)
    for t in range(n_types):
        sel = P["shape"] == t
        S = shapes[t]
        s = P["scale"][sel]
        P["inv_mass"][sel] = 1.0 / (S["mass"] * s ** 3)
        P["inv_inertia"][sel] = S["inv_inertia"][None, :] / (s ** 5)[:, None]
    P["kn"][:] = kn
    P["ks"][:] = 0.4 * kn
    P["mu"][:] = mu
    P["e"][:] = restitution
    zmax = top + 0.05
    # y range chosen so that max.y >= max.z: the reference tests ghost z against max.y
    y0 = zmax - cell_xy
    P["pos"][:, 1] += y0
    cfg = {"domain_min": (0.0, y0, 0.0), "domain_max": (cell_xy, y0 + cell_xy, zmax), "gravity": (0.0, 0.0, -gravity),
           "dt": dt, "periodic_x": 1, "periodic_y": 1, "quirks": quirks}
    sc = Scene(cfg=cfg, shapes=shapes, particles=P, meta={"name": "C2", "seed": seed, "bed_top": top})
    sc.walls = floor_wall(shapes, (0.5 * cell_xy, y0 + 0.5 * cell_xy), 0.0, 0.75 * cell_xy, mu, restitution)
    return sc


def _relocate_on_top(sc: Scene, idx: np.ndarray, rng) -> None:
    """Move particles `idx` into loose layers (pitch = 1.02 x their largest bounding diameter)
    above the bed, then re-derive the domain (z top, and the y range so that max.y >= max.z)."""
    P = sc.particles
    cell_xy = sc.cfg["domain_max"][0] - sc.cfg["domain_min"][0]
    y0 = sc.cfg["domain_min"][1]
    rb = np.array([np.linalg.norm(s["nodes"], axis=1).max() if len(s["nodes"]) else 0.0
                   for s in sc.shapes])[P["shape"]] * P["scale"]
    keep = np.ones(len(rb), bool)
    keep[idx] = False
    z = float((P["pos"][keep, 2] + rb[keep]).max()) if keep.any() else 0.0
    k = 0
    while k < len(idx):
        pitch = 2.0 * rb[idx].max() * 1.02
        m = max(1, int(np.floor(cell_xy / pitch)))
        take = idx[k:k + m * m]
        j = np.arange(len(take))
        P["pos"][take, 0] = (j % m + 0.5) * (cell_xy / m)
        P["pos"][take, 1] = y0 + (j // m + 0.5) * (cell_xy / m)
        P["pos"][take, 2] = z + 0.5 * pitch
        z += pitch
        k += len(take)
    top = float((P["pos"][:, 2] + rb).max())
    zmax = top + 0.05
    new_y0 = zmax - cell_xy
    P["pos"][:, 1] += new_y0 - y0
    if sc.walls and len(sc.walls["shape"]):
        sc.walls["pos"][:, 1] += new_y0 - y0
    sc.cfg["domain_min"] = (sc.cfg["domain_min"][0], new_y0, sc.cfg["domain_min"][2])
    sc.cfg["domain_max"] = (sc.cfg["domain_max"][0], new_y0 + cell_xy, zmax)
    sc.meta["bed_top"] = top


def remove_initial_overlaps(make_sim, sc: Scene, seed: int = 7, rounds: int = 12, tol: float = 5e-6,
                            log=None) -> int:
    """Random-lattice start without significant initial overlaps: resample the orientation of
    every particle overlapping a neighbour, a wall or a periodic image by more than `tol`;
    particles still overlapping after `rounds` rounds are moved into loose layers on top of the
    bed (the particle count is unchanged).  `make_sim(cfg)` returns an un-loaded simulation
    context; detection runs with gravity off and a vanishing time step.  Returns the number of
    relocated particles."""
    rng = np.random.default_rng(seed)
    P = sc.particles
    n = len(P["shape"])
    node0 = np.concatenate([[0], np.cumsum([len(sc.shapes[s]["nodes"]) for s in P["shape"]])])
    zero = np.zeros((n, 3))

    def overlapping(sim):
        sim.set_particles(pos=P["pos"], quat=P["quat"], vel=zero, angvel=zero, force=zero, torque=zero)
        sim.step(1)
        bad = set()
        for lst in range(5):
            c = sim.contacts(lst)
            deep = c["overlap"] > tol
            if deep.any():
                bad.update((np.searchsorted(node0, c["master"][deep], side="right") - 1).tolist())
        return np.array(sorted(bad), dtype=np.int64)

    def detector():
        cfg = dict(sc.cfg)
        cfg["gravity"] = (0.0, 0.0, 0.0)
        cfg["dt"] = 1e-30
        return sc.load(make_sim(cfg))

    sim = detector()
    for it in range(rounds):
        bad = overlapping(sim)
        if log:
            log(f"[scene] overlap round {it}: {len(bad)} particles overlap")
        if len(bad) == 0:
            sim.close()
            return 0
        P["quat"][bad] = random_quaternions(rng, len(bad))
    bad = overlapping(sim)
    sim.close()
    moved = 0
    while len(bad):
        moved += len(bad)
        _relocate_on_top(sc, bad, rng)
        sim = detector()
        bad = overlapping(sim)
        sim.close()
        if log:
            log(f"[scene] relocated {moved} particles; {len(bad)} still overlap")
    return moved


# ----------------------------------------------------------------------------------------------
# Config 3: bonded agglomerate of 3 superellipsoid types, compressed between two walls
# ----------------------------------------------------------------------------------------------
def quat_rotate(q: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Rotate vectors v (..., 3) by unit quaternions q (..., 4), scalar first."""
    w, u = q[..., :1], q[..., 1:]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def sdf_trilinear(shape: dict, p: np.ndarray) -> np.ndarray:
    """Trilinear value of a shape's grid at unscaled local points p (n, 3); +inf outside the grid."""
    sdf = shape["sdf"].astype(np.float64)
    nz, ny, nx = sdf.shape
    g = (p - shape["origin"]) / shape["spacing"]
    i0 = np.floor(g).astype(np.int64)
    ok = np.all((i0 >= 0) & (i0 < np.array([nx - 1, ny - 1, nz - 1])), axis=1)
    i0 = np.clip(i0, 0, np.array([nx - 2, ny - 2, nz - 2]))
    f = g - i0
    out = np.zeros(len(p))
    for c in range(8):
        dx, dy, dz = c & 1, (c >> 1) & 1, c >> 2
        w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * (f[:, 2] if dz else 1 - f[:, 2])
        out += w * sdf[i0[:, 2] + dz, i0[:, 1] + dy, i0[:, 0] + dx]
    out[~ok] = np.inf
    return out


def node_owner(sc: Scene, nodes: np.ndarray) -> np.ndarray:
    """Particle owning each boundary node index (nodes are numbered particle by particle)."""
    node0 = np.concatenate([[0], np.cumsum([len(sc.shapes[s]["nodes"]) for s in sc.particles["shape"]])])
    return np.searchsorted(node0, nodes, side="right") - 1


def macro_stats(sc: Scene, parts: dict, lists: list, z_band=(0.1, 0.6)) -> dict:
    """Macroscopic state of a simulation, the quantities long runs are compared on (trajectories
    of DEM diverge chaotically, these do not):
    - solid fraction: volume of the particles whose centres lie in the height band
      [z_band[0] H, z_band[1] H] above the domain floor, over the band's volume; H is the 95th
      percentile of the centre heights and the band spans the domain's x-y extent;
    - coordination number: distinct contacting particle pairs (particle-particle and periodic-image
      lists), counted once for each of the two particles, per particle; wall contacts apart;
    - kinetic energy: translational 1/2 m v^2 plus rotational 1/2 w.(R I R^T) w.
    `parts` is get_particles(); `lists` the five contact lists (contacts(l) for l in 0..4)."""
    P = sc.particles
    cfg = sc.cfg
    lo, hi = np.asarray(cfg["domain_min"], float), np.asarray(cfg["domain_max"], float)
    m = 1.0 / P["inv_mass"]
    vol = m / np.array([sc.shapes[t]["density"] for t in P["shape"]])
    z = parts["pos"][:, 2] - lo[2]
    H = float(np.percentile(z, 95))
    z0, z1 = z_band[0] * H, z_band[1] * H
    inband = (z >= z0) & (z < z1)
    area = (hi[0] - lo[0]) * (hi[1] - lo[1])
    phi = float(vol[inband].sum() / (area * (z1 - z0))) if z1 > z0 else 0.0
    n = len(m)
    pairs = set()
    for l in (0, 2, 3, 4):
        c = lists[l]
        if len(c["master"]) == 0:
            continue
        a = node_owner(sc, c["master"])
        b = c["slave"].astype(np.int64)
        pairs.update(zip(np.minimum(a, b).tolist(), np.maximum(a, b).tolist()))
    z_pp = 2.0 * sum(1 for a, b in pairs if a != b) / n
    wall = set()
    c = lists[1]
    if len(c["master"]):
        wall.update(zip(node_owner(sc, c["master"]).tolist(), c["slave"].tolist()))
    v, w, q = parts["vel"], parts["angvel"], parts["quat"]
    ke_t = 0.5 * float(np.sum(m * np.sum(v * v, axis=1)))
    Ii = P["inv_inertia"]
    I = np.array([inv_sym(x) for x in Ii])  # local inertia (xx, yy, zz, xy, xz, yz)
    wl = quat_rotate(np.concatenate([q[:, :1], -q[:, 1:]], axis=1), w)  # body frame
    Iw = np.stack([I[:, 0] * wl[:, 0] + I[:, 3] * wl[:, 1] + I[:, 4] * wl[:, 2],
                   I[:, 3] * wl[:, 0] + I[:, 1] * wl[:, 1] + I[:, 5] * wl[:, 2],
                   I[:, 4] * wl[:, 0] + I[:, 5] * wl[:, 1] + I[:, 2] * wl[:, 2]], axis=1)
    ke_r = 0.5 * float(np.sum(wl * Iw))
    return {"solid_fraction": phi, "coordination": z_pp, "wall_contacts_per_particle": len(wall) / n,
            "kinetic_energy": ke_t + ke_r, "ke_translational": ke_t, "ke_rotational": ke_r,
            "bed_height": H, "band": (z0, z1)}


def surface_gap(sc: Scene, i: int, j: int):
    """Smallest signed distance from particle i's boundary nodes to particle j's surface (world
    units), and the node position (world) where it occurs."""
    P, S = sc.particles, sc.shapes
    si, sj = S[P["shape"][i]], S[P["shape"][j]]
    xw = quat_rotate(P["quat"][i][None, :], si["nodes"] * P["scale"][i]) + P["pos"][i]
    loc = quat_rotate(np.array([P["quat"][j][0], *(-P["quat"][j][1:])])[None, :], xw - P["pos"][j]) / P["scale"][j]
    phi = sdf_trilinear(sj, loc) * P["scale"][j]
    k = int(np.argmin(phi))
    return float(phi[k]), xw[k]


def bond_pairs_by_gap(sc: Scene, max_gap: float, radius_mult: float, youngs: float, poisson: float, rng,
                      weibull_shape: float = 5.0, weibull_scale: float = 1e6, friction: float = 0.0):
    """One parallel bond per particle pair whose surfaces are closer than max_gap (both directions
    tested): bond point midway across the gap, normal from the slave's centre towards the master's,
    length = centre distance, radius = radius_mult x the smaller equivalent radius, tensile strength
    and cohesion Weibull(shape, scale)."""
    P = sc.particles
    n = sc.n_particles
    rb = np.array([np.linalg.norm(s["nodes"], axis=1).max() if len(s["nodes"]) else 0.0 for s in sc.shapes])
    rbound = rb[P["shape"]] * P["scale"]
    x = P["pos"]
    bonds = []
    for i in range(n):
        d = np.linalg.norm(x[i + 1:] - x[i], axis=1)
        for jj in np.nonzero(d <= rbound[i] + rbound[i + 1:] + max_gap)[0]:
            j = i + 1 + int(jj)
            g1, p1 = surface_gap(sc, i, j)
            g2, p2 = surface_gap(sc, j, i)
            g, p = (g1, p1) if g1 <= g2 else (g2, p2)
            if not (g < max_gap):
                continue
            nrm = x[i] - x[j]
            L = float(np.linalg.norm(nrm))
            nrm = nrm / L
            sgn = 1.0 if g1 <= g2 else -1.0  # node of i: point moves towards j, node of j: towards i
            point = p - sgn * 0.5 * max(g, 0.0) * nrm
            r = radius_mult * min(P["scale"][i], P["scale"][j])
            st, co = weibull_scale * rng.weibull(weibull_shape, 2)
            bonds.append((i, j, tuple(point), tuple(nrm), r, L, youngs, poisson, float(st), float(co), friction))
    return bonds


def scene_c3(n: int = 100_000, seed: int = 20261205, n_grid: int = 40, n_nodes: int = 400,
             exponents=(2.0, 3.0, 6.0), r_range=(2e-3, 3e-3), pitch: float | None = None, jitter: float = 0.05,
             dt: float = 2e-6, kn: float = 1e5, restitution: float = 0.5, mu: float = 0.5, density: float = 2500.0,
             gravity: float = 0.0, wall_speed: float = 0.05, bond_gap: float = 1e-4, radius_mult: float = 0.8,
             youngs: float = 1e8, poisson: float = 0.3, weibull_shape: float = 5.0, weibull_scale: float = 1e6,
             quirks: int = 7, bonds: bool = True) -> Scene:
    """Lattice of volume-equivalent radii U(r_range) in a cube, random orientations; bonds between
    pairs closer than bond_gap; top and bottom walls close at wall_speed each."""
    rng = np.random.default_rng(seed)
    shapes: list = []
    rb = []
    for p in exponents:
        e = 2.0 / p
        ax = np.ones(3)
        g = np.linspace(-1, 1, 121)
        Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
        f, _ = superellipsoid_value_grad(np.stack([X, Y, Z], -1), 1.0, 1.0, 1.0, e, e)
        vol = (f <= 0).mean() * 8.0
        ax = ax * (4.0 / 3.0 * np.pi / vol) ** (1.0 / 3.0)
        h = 2.0 * ax.max() / (n_grid - 5)
        sdf, org = superellipsoid_grid(*ax, e, e, n_grid, h)
        nodes = superellipsoid_nodes(*ax, e, e, n_nodes)
        sid = _add_shape(shapes, sdf, org, h, nodes, density)
        rb.append(np.linalg.norm(shapes[sid]["nodes"], axis=1).max())
    rb = np.asarray(rb)
    P = _particles(n)
    P["shape"] = rng.integers(0, len(exponents), n).astype(np.int32)
    P["scale"] = rng.uniform(r_range[0], r_range[1], n)
    if pitch is None:
        pitch = 2.0 * r_range[1] * 1.005
    m = int(np.ceil(n ** (1.0 / 3.0)))
    k = np.arange(n)
    idx = np.stack([k % m, (k // m) % m, k // (m * m)], 1).astype(float)
    P["pos"] = (idx + 0.5) * pitch + rng.uniform(-jitter, jitter, (n, 3)) * pitch * 0.1
    P["quat"] = random_quaternions(rng, n)
    for t in range(len(exponents)):
        sel = P["shape"] == t
        Sh = shapes[t]
        s = P["scale"][sel]
        P["inv_mass"][sel] = 1.0 / (Sh["mass"] * s ** 3)
        P["inv_inertia"][sel] = Sh["inv_inertia"][None, :] / (s ** 5)[:, None]
    P["kn"][:] = kn
    P["ks"][:] = 0.4 * kn
    P["mu"][:] = mu
    P["e"][:] = restitution
    side = m * pitch
    lo, hi = np.zeros(3), np.array([side, side, (np.max(idx[:, 2]) + 1) * pitch])
    margin = 0.2 * side + 0.01
    cfg = {"domain_min": tuple(lo - margin), "domain_max": tuple(hi + margin), "gravity": (0.0, 0.0, -gravity),
           "dt": dt, "quirks": quirks}
    sc = Scene(cfg=cfg, shapes=shapes, particles=P, meta={"name": "C3", "seed": seed})
    sc.walls = box_walls(shapes, lo, hi, mu, restitution)
    sc.walls["vel"][0] = (0.0, 0.0, wall_speed)    # bottom wall moves up
    sc.walls["vel"][1] = (0.0, 0.0, -wall_speed)   # top wall moves down
    if bonds:
        sc.bonds = bond_pairs_by_gap(sc, bond_gap, radius_mult, youngs, poisson, rng, weibull_shape, weibull_scale)
    return sc
