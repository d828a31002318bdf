"""ctypes binding of the LS-DEM C ABI (include/lsdem.h).

`Sim` drives one context of a library that exports the lsdem_* entry points.  The product
library is liblsdem.so next to this file; `load_library()` raises if it has not been built
(there is no CPU fallback).  The same class accepts any library exporting the same entry
points under another prefix, which is how the tests drive the host restatement.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liblsdem.so"

LSDEM_OK = 0
LSDEM_ERR_ARG = 1
LSDEM_ERR_STATE = 2
LSDEM_ERR_HIP = 3
LSDEM_ERR_CAPACITY = 4
QUIRK_INT_X_GRADIENT = 1
QUIRK_GHOST_Z_VS_MAXY = 2
QUIRK_WRAP_AFTER_BUILD = 4
QUIRK_ALL = 7
LIST_PP, LIST_PW, LIST_GX, LIST_GY, LIST_GXY = range(5)


class Config(C.Structure):
    _fields_ = [
        ("domain_min", C.c_double * 3),
        ("domain_max", C.c_double * 3),
        ("gravity", C.c_double * 3),
        ("dt", C.c_double),
        ("periodic_x", C.c_int32),
        ("periodic_y", C.c_int32),
        ("quirks", C.c_int32),
        ("block_size", C.c_int32),
        ("pair_capacity", C.c_int64),
        ("contact_capacity", C.c_int64),
        ("device", C.c_int32),
        ("no_sdf_cull", C.c_int32),
        ("skin", C.c_double),
        ("no_certificates", C.c_int32),
        ("broad_phase_launches", C.c_int32),
        ("broad_phase_blocks", C.c_int32),
        ("narrow_batch", C.c_int32),
        ("no_fused_steps", C.c_int32),
        ("no_recovery", C.c_int32),
        ("barrier_spin", C.c_int32),
        ("broad_phase_coop", C.c_int32),
        ("periodic_sector", C.c_int32),
        ("narrow_blocks", C.c_int32),
    ]


class SurrogateParams(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("batch", C.c_int32), ("samples", C.c_int64), ("validation", C.c_int64),
        ("lr", C.c_double), ("lr_final", C.c_double), ("clamp", C.c_double), ("band_factor", C.c_double),
        ("audit_every", C.c_int32), ("pad", C.c_int32), ("seed", C.c_uint64), ("reserved", C.c_int64 * 8),
    ]


class SurrogateReport(C.Structure):
    _fields_ = [
        ("trained", C.c_int32), ("enabled", C.c_int32), ("n_shapes", C.c_int32), ("n_params", C.c_int32),
        ("iterations", C.c_int64), ("final_loss", C.c_double), ("max_error", C.c_double),
        ("mean_error", C.c_double), ("band", C.c_double), ("train_ms", C.c_double), ("length_scale", C.c_double),
        ("audits", C.c_int64), ("audited_contacts", C.c_int64), ("missed_contacts", C.c_int64),
        ("reserved", C.c_int64 * 8),
    ]


SURROGATE_N_PARAM = 19777


class Counters(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("pairs", C.c_int64),
        ("node_tests", C.c_int64),
        ("sdf_queries", C.c_int64),
        ("contacts", C.c_int64),
        ("pair_capacity", C.c_int64),
        ("contact_capacity", C.c_int64),
        ("recoveries", C.c_int64),
        ("reserved", C.c_int64 * 8),
    ]


NUM_PHASES = 8
PHASES = ("integrate", "sort", "pairs", "reaction", "pair_check", "candidates", "bonds", "reduce")


class Profile(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("ms", C.c_double * NUM_PHASES),
        ("pairs", C.c_int64),
        ("node_tests", C.c_int64),
        ("sdf_queries", C.c_int64),
        ("contacts", C.c_int64),
        ("cluster_tests", C.c_int64),
        ("master_nodes", C.c_int64),
        ("evaluated_pairs", C.c_int64),
        ("rebuilds", C.c_int64),
        ("sampled_steps", C.c_int64),
        ("reserved", C.c_int64 * 3),
    ]


class DeviceView(C.Structure):
    """lsdem_device_view: N x 3 AoS arrays (device pointers on the GPU, host pointers in the oracle)."""
    _fields_ = [
        ("force", C.c_void_p),
        ("torque", C.c_void_p),
        ("velocity", C.c_void_p),
        ("angular_velocity", C.c_void_p),
        ("position", C.c_void_p),
        ("quaternion", C.c_void_p),
        ("num", C.c_int64),
        ("grid_dim", C.c_int64),
        ("block_dim", C.c_int64),
        ("stream", C.c_void_p),
    ]


class WallDeviceView(C.Structure):
    """lsdem_wall_device_view: W x 3 (W x 4 for the orientation) device arrays of the walls."""
    _fields_ = [
        ("position", C.c_void_p),
        ("quaternion", C.c_void_p),
        ("velocity", C.c_void_p),
        ("angular_velocity", C.c_void_p),
        ("num", C.c_int64),
        ("grid_dim", C.c_int64),
        ("block_dim", C.c_int64),
        ("stream", C.c_void_p),
    ]


# void (*)(void* user, double time, const lsdem_device_view* view)
EXTERNAL_FORCE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.POINTER(DeviceView))


class LSDEMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lsdem error {code}: {msg}")
        self.code = code


_PD = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _p(a):
    """Pointer of a contiguous numpy array (or NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


_libs: dict = {}


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load liblsdem.so.  Raises FileNotFoundError when it has not been built."""
    path = Path(path) if path else LIB_PATH
    if not path.exists():
        raise FileNotFoundError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    key = str(path)
    if key not in _libs:
        _libs[key] = C.CDLL(key)
    return _libs[key]


class Sim:
    """One simulation context.  `prefix` selects the entry-point family of `lib`."""

    def __init__(self, cfg: dict, lib: C.CDLL | None = None, prefix: str = "lsdem_"):
        self.lib = lib if lib is not None else load_library()
        self.prefix = prefix
        c = Config()
        c.domain_min[:] = cfg.get("domain_min", (0.0, 0.0, 0.0))
        c.domain_max[:] = cfg.get("domain_max", (1.0, 1.0, 1.0))
        c.gravity[:] = cfg.get("gravity", (0.0, 0.0, 0.0))
        c.dt = cfg.get("dt", 1e-5)
        c.periodic_x = int(cfg.get("periodic_x", 0))
        c.periodic_y = int(cfg.get("periodic_y", 0))
        c.quirks = int(cfg.get("quirks", QUIRK_ALL))
        c.block_size = int(cfg.get("block_size", 256))
        c.pair_capacity = int(cfg.get("pair_capacity", 0))
        c.contact_capacity = int(cfg.get("contact_capacity", 0))
        c.device = int(cfg.get("device", 0))
        c.no_sdf_cull = int(cfg.get("no_sdf_cull", 0))
        c.no_certificates = int(cfg.get("no_certificates", 0))
        c.broad_phase_launches = int(cfg.get("broad_phase_launches", 0))
        c.broad_phase_blocks = int(cfg.get("broad_phase_blocks", 0))
        c.narrow_batch = int(cfg.get("narrow_batch", 0))
        c.no_fused_steps = int(cfg.get("no_fused_steps", 0))
        c.no_recovery = int(cfg.get("no_recovery", 0))
        c.barrier_spin = int(cfg.get("barrier_spin", 0))
        c.broad_phase_coop = int(cfg.get("broad_phase_coop", 0))
        c.periodic_sector = int(cfg.get("periodic_sector", 0))
        c.narrow_blocks = int(cfg.get("narrow_blocks", 0))
        c.skin = float(cfg.get("skin", -1.0))
        self.cfg = c
        self.ctx = C.c_void_p()
        self._call("create", C.byref(c), C.byref(self.ctx))
        self.n_particles = 0
        self.n_walls = 0
        self.n_bonds = 0

    # -- plumbing --
    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _err(self) -> str:
        f = self._fn("last_error")
        f.restype = C.c_char_p
        s = f()
        return s.decode() if s else ""

    def _call(self, name, *args):
        f = self._fn(name)
        f.restype = C.c_int
        rc = f(*args)
        if rc != LSDEM_OK:
            raise LSDEMError(rc, f"{self.prefix}{name}: {self._err()}")
        return rc

    def close(self):
        if self.ctx:
            f = self._fn("destroy")
            f.restype = None
            f(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup --
    def add_shape(self, sdf: np.ndarray, origin, spacing: float, nodes: np.ndarray) -> int:
        """sdf: float32 array of shape (nz, ny, nx) (x fastest); nodes: (m, 3) float64."""
        sdf = np.ascontiguousarray(sdf, dtype=np.float32)
        nz, ny, nx = sdf.shape
        org = (C.c_double * 3)(*map(float, origin))
        nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 3)
        sid = C.c_int32()
        self._call("add_shape", self.ctx, _p(sdf), C.c_int32(nx), C.c_int32(ny), C.c_int32(nz), org, C.c_double(spacing),
                   _p(nodes), C.c_int32(len(nodes)), C.byref(sid))
        return sid.value

    def add_particles(self, shape, scale, pos, quat, vel, angvel, inv_mass, inv_inertia, kn, ks, mu, e):
        n = len(shape)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (scale, pos, quat, vel, angvel, inv_mass,
                                                                 inv_inertia, kn, ks, mu, e)]
        sh = np.ascontiguousarray(shape, dtype=np.int32)
        self._call("add_particles", self.ctx, C.c_int32(n), _p(sh), *[_p(x) for x in a])
        self.n_particles += n

    def add_walls(self, shape, radius, pos, quat, vel, angvel, mu, e):
        n = len(shape)
        sh = np.ascontiguousarray(shape, dtype=np.int32)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (radius, pos, quat, vel, angvel, mu, e)]
        self._call("add_walls", self.ctx, C.c_int32(n), _p(sh), *[_p(x) for x in a])
        self.n_walls += n

    def add_bond(self, i, j, point, normal, radius, length, youngs, poisson, tensile=0.0, cohesion=0.0,
                 friction=0.0):
        pt = (C.c_double * 3)(*map(float, point))
        nm = (C.c_double * 3)(*map(float, normal))
        self._call("add_bond", self.ctx, C.c_int32(i), C.c_int32(j), pt, nm, C.c_double(radius), C.c_double(length),
                   C.c_double(youngs), C.c_double(poisson), C.c_double(tensile), C.c_double(cohesion),
                   C.c_double(friction))
        self.n_bonds += 1

    def initialize(self):
        self._call("initialize", self.ctx)

    # -- stepping and state --
    def step(self, n: int = 1):
        self._call("step", self.ctx, C.c_int32(n))

    def begin_phase(self):
        """The pre-loop build of a solve phase (LSSolver.h:996-998, 1061-1066): contacts at the
        current poses, periodic wrap, ghost images; no step is taken.  Runs by itself before the
        first step after initialize()."""
        self._call("begin_phase", self.ctx)

    def step_graph(self, n: int):
        """n steps (even) replayed from a HIP graph captured on the first call with that count."""
        self._call("step_graph", self.ctx, C.c_int32(n))

    STEP_PHASES = ("integrate", "hash", "pairs", "narrow", "forces", "reduce")

    def phase(self, name: str):
        """One phase of a step driven phase by phase (GPU library only): integrate, hash, pairs,
        narrow, forces, reduce, in that order; the six make one step."""
        if name not in self.STEP_PHASES:
            raise ValueError(f"unknown phase {name!r}")
        self._call("phase_" + name, self.ctx)

    def cell_table(self) -> dict:
        """Search grid and cell table of the last broad phase (GPU library only)."""
        org = (C.c_double * 3)()
        edge = C.c_double()
        dims = (C.c_int32 * 3)()
        self._call("get_cell_table", self.ctx, org, C.byref(edge), dims, None, None, C.c_int64(0), None)
        n = self.n_particles
        ncell = int(dims[0]) * int(dims[1]) * int(dims[2])
        cell_of = np.zeros(n, np.int32)
        order = np.zeros(n, np.int32)
        start = np.zeros(ncell + 1, np.int32)
        self._call("get_cell_table", self.ctx, None, None, None, _p(cell_of), _p(order),
                   C.c_int64(ncell + 1), _p(start))
        return {"origin": np.array(org[:]), "cell": edge.value, "dims": np.array(dims[:]), "cell_of": cell_of,
                "sorted": order, "cell_start": start}

    def pairs(self) -> dict:
        """Candidate pairs of the last broad phase (GPU library only): master, partner, type."""
        cnt = C.c_int64()
        self._call("get_pairs", self.ctx, C.c_int64(0), C.byref(cnt), None, None, None)
        m = cnt.value
        out = {k: np.zeros(m, np.int32) for k in ("master", "partner", "type")}
        if m:
            self._call("get_pairs", self.ctx, C.c_int64(m), C.byref(cnt), _p(out["master"]),
                       _p(out["partner"]), _p(out["type"]))
        return out

    def set_external_force(self, fn):
        """fn(time: float, view: DeviceView) runs after the contact and bond forces of every step and
        before the second half-kick; None removes it."""
        if fn is None:
            self._hook = None
            self._call("set_external_force", self.ctx, EXTERNAL_FORCE_FN(), None)
            return
        self._hook = EXTERNAL_FORCE_FN(lambda user, time, view: fn(time, view.contents))
        self._call("set_external_force", self.ctx, self._hook, None)

    def wall_device_view(self) -> WallDeviceView:
        v = WallDeviceView()
        self._call("get_wall_device_view", self.ctx, C.byref(v))
        return v

    def device_view(self) -> DeviceView:
        v = DeviceView()
        self._call("get_device_view", self.ctx, C.byref(v))
        return v

    # -- learned culling surrogate (GPU library only) --
    def surrogate_train(self, **kw) -> dict:
        p = SurrogateParams()
        self.lib.lsdem_surrogate_params_default(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        self._call("surrogate_train", self.ctx, C.byref(p))
        return self.surrogate_report()

    def surrogate_enable(self, on: bool = True):
        self._call("surrogate_enable", self.ctx, C.c_int32(int(on)))

    def surrogate_set_band(self, band: float, audit_every: int):
        self._call("surrogate_set_band", self.ctx, C.c_double(band), C.c_int32(audit_every))

    def surrogate_report(self) -> dict:
        r = SurrogateReport()
        self._call("surrogate_get_report", self.ctx, C.byref(r))
        return {k: getattr(r, k) for k, _ in SurrogateReport._fields_ if k != "reserved"}

    def surrogate_weights(self, n_shapes: int) -> np.ndarray:
        w = np.zeros(SURROGATE_N_PARAM + 16 * n_shapes, np.float32)
        self._call("surrogate_get_weights", self.ctx, w.ctypes.data_as(C.c_void_p), C.c_int64(w.size))
        return w

    def surrogate_set_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, np.float32)
        self._call("surrogate_set_weights", self.ctx, w.ctypes.data_as(C.c_void_p), C.c_int64(w.size))

    def surrogate_eval(self, points, shape_ids) -> np.ndarray:
        p = np.ascontiguousarray(points, np.float32)
        s = np.ascontiguousarray(shape_ids, np.int32)
        out = np.zeros(len(s), np.float32)
        self._call("surrogate_eval", self.ctx, C.c_int64(len(s)), p.ctypes.data_as(C.c_void_p),
                   s.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
        return out

    def surrogate_gradient(self, points, shape_ids, targets, n_shapes: int):
        p = np.ascontiguousarray(points, np.float32)
        s = np.ascontiguousarray(shape_ids, np.int32)
        t = np.ascontiguousarray(targets, np.float32)
        g = np.zeros(SURROGATE_N_PARAM + 16 * n_shapes, np.float32)
        loss = C.c_double()
        self._call("surrogate_gradient", self.ctx, C.c_int64(len(s)), p.ctypes.data_as(C.c_void_p),
                   s.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p),
                   C.byref(loss))
        return loss.value, g

    def time(self) -> float:
        t = C.c_double()
        self._call("get_time", self.ctx, C.byref(t))
        return t.value

    def get_particles(self) -> dict:
        n = self.n_particles
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "angvel", "force", "torque")}
        out["quat"] = np.zeros((n, 4))
        self._call("get_particles", self.ctx, _p(out["pos"]), _p(out["quat"]), _p(out["vel"]), _p(out["angvel"]),
                   _p(out["force"]), _p(out["torque"]))
        return out

    def set_particles(self, pos=None, quat=None, vel=None, angvel=None, force=None, torque=None):
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float64)
             for x in (pos, quat, vel, angvel, force, torque)]
        self._call("set_particles", self.ctx, *[_p(x) for x in a])

    def get_walls(self) -> dict:
        n = self.n_walls
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "angvel")}
        out["quat"] = np.zeros((n, 4))
        self._call("get_walls", self.ctx, _p(out["pos"]), _p(out["quat"]), _p(out["vel"]), _p(out["angvel"]))
        return out

    def set_walls(self, pos=None, quat=None, vel=None, angvel=None):
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (pos, quat, vel, angvel)]
        self._call("set_walls", self.ctx, *[_p(x) for x in a])

    def contacts(self, lst: int) -> dict:
        """Contacts of one list of the last step, sorted by (master node, slave)."""
        cnt = C.c_int64()
        self._call("get_contacts", self.ctx, C.c_int32(lst), C.c_int64(0), C.byref(cnt),
                   None, None, None, None, None, None, None, None)
        n = cnt.value
        out = {
            "master": np.zeros(n, np.int32), "slave": np.zeros(n, np.int32),
            "point": np.zeros((n, 3)), "normal": np.zeros((n, 3)), "overlap": np.zeros(n),
            "spring": np.zeros((n, 3)), "energy_n": np.zeros(n), "energy_s": np.zeros(n),
        }
        self._call("get_contacts", self.ctx, C.c_int32(lst), C.c_int64(n), C.byref(cnt),
                   _p(out["master"]), _p(out["slave"]), _p(out["point"]), _p(out["normal"]), _p(out["overlap"]),
                   _p(out["spring"]), _p(out["energy_n"]), _p(out["energy_s"]))
        return out

    def bonds(self) -> dict:
        n = self.n_bonds
        out = {"active": np.zeros(n, np.int32), "point": np.zeros((n, 3)), "energies": np.zeros((n, 4)),
               "sigma_max": np.zeros(n), "tau_max": np.zeros(n)}
        if n:
            self._call("get_bonds", self.ctx, _p(out["active"]), _p(out["point"]), _p(out["energies"]),
                       _p(out["sigma_max"]), _p(out["tau_max"]))
        return out

    def profile_begin(self, max_steps: int, every: int = 1):
        """Time every `every`-th step's phases, at most max_steps of them; counters cover all steps."""
        if every == 1:
            self._call("profile_begin", self.ctx, C.c_int32(max_steps))
        else:
            self._call("profile_begin_sampled", self.ctx, C.c_int32(max_steps), C.c_int32(every))

    def profile_end(self) -> dict:
        p = Profile()
        self._call("profile_end", self.ctx, C.byref(p))
        return {"steps": p.steps, "sampled_steps": p.sampled_steps, "ms": dict(zip(PHASES, list(p.ms))),
                "pairs": p.pairs, "node_tests": p.node_tests,
                "sdf_queries": p.sdf_queries, "contacts": p.contacts, "cluster_tests": p.cluster_tests,
                "master_nodes": p.master_nodes, "evaluated_pairs": p.evaluated_pairs, "rebuilds": p.rebuilds}

    def counters(self) -> dict:
        c = Counters()
        self._call("get_counters", self.ctx, C.byref(c))
        return {k: getattr(c, k) for k, _ in Counters._fields_ if k != "reserved"}


def mass_properties(lib, sdf: np.ndarray, origin, spacing: float, density: float, prefix: str = "lsdem_"):
    """Mass, centroid, inertia and inverse inertia of an SDF grid (smoothed Heaviside)."""
    sdf = np.ascontiguousarray(sdf, dtype=np.float64)
    nz, ny, nx = sdf.shape
    org = (C.c_double * 3)(*map(float, origin))
    m = C.c_double()
    cen = (C.c_double * 3)()
    I = np.zeros(6)
    Ii = np.zeros(6)
    f = getattr(lib, prefix + "mass_properties")
    f.restype = C.c_int
    rc = f(_p(sdf), C.c_int32(nx), C.c_int32(ny), C.c_int32(nz), org, C.c_double(spacing), C.c_double(density),
           C.byref(m), cen, _p(I), _p(Ii))
    if rc != 0:
        raise LSDEMError(rc, "mass_properties failed")
    return m.value, np.array(cen[:]), I, Ii
