"""Helpers comparing two simulations (GPU library vs oracle) list by list."""
from __future__ import annotations

import numpy as np

LISTS = {0: "pp", 1: "pw", 2: "gx", 3: "gy", 4: "gxy"}


def compare_contacts(a: dict, b: dict, name: str, tol: float = 1e-12):
    """Bit-exact (master, slave) sets; overlap/point/normal within tol relative; spring within 1e-9."""
    assert len(a["master"]) == len(b["master"]), f"{name}: {len(a['master'])} vs {len(b['master'])} contacts"
    assert np.array_equal(a["master"], b["master"]), f"{name}: master ids differ"
    assert np.array_equal(a["slave"], b["slave"]), f"{name}: slave ids differ"
    if len(a["master"]) == 0:
        return
    ov = np.abs(a["overlap"] - b["overlap"]).max() / (np.abs(b["overlap"]).max() + 1e-300)
    assert ov <= tol, f"{name}: overlap rel err {ov}"
    for k in ("point", "normal"):
        e = np.abs(a[k] - b[k]).max() / (np.abs(b[k]).max() + 1e-300)
        assert e <= tol, f"{name}: {k} rel err {e}"
    sc = np.abs(b["spring"]).max()
    if sc > 0:
        e = np.abs(a["spring"] - b["spring"]).max() / sc
        assert e <= 1e-9, f"{name}: spring rel err {e}"


def compare_state(a: dict, b: dict, ftol: float = 1e-5):
    """Forces/torques within ftol relative to the largest body force/torque of the step."""
    out = {}
    for k in ("force", "torque"):
        sc = np.abs(b[k]).max()
        e = np.abs(a[k] - b[k]).max() / sc if sc > 0 else np.abs(a[k]).max()
        assert e <= ftol, f"{k} rel err {e}"
        out[k] = e
    for k in ("pos", "quat", "vel", "angvel"):
        sc = np.abs(b[k]).max() + 1e-300
        out[k] = np.abs(a[k] - b[k]).max() / sc
    return out


def compare_bonds(a: dict, b: dict, name: str, tol: float = 1e-9):
    """Identical active flags (every break event on the same step); energies, peak stresses and
    bond points within tol relative to the largest value."""
    assert np.array_equal(a["active"], b["active"]), f"{name}: bond states differ"
    for k in ("energies", "sigma_max", "tau_max", "point"):
        sc = np.abs(b[k]).max() if b[k].size else 0.0
        if sc > 0:
            e = np.abs(a[k] - b[k]).max() / sc
            assert e <= tol, f"{name}: bond {k} rel err {e}"


def step_and_compare(gpu, ref, steps: int, lists=(0, 1), ftol: float = 1e-5, log=None, bonds: bool = False,
                     ctol: float = 1e-12):
    worst = {}
    counts = []
    for s in range(steps):
        gpu.step(1)
        ref.step(1)
        row = {}
        if bonds:
            ba, bb = gpu.bonds(), ref.bonds()
            compare_bonds(ba, bb, f"step {s}")
            row["broken"] = int((bb["active"] == 0).sum())
        for l in lists:
            ca, cb = gpu.contacts(l), ref.contacts(l)
            compare_contacts(ca, cb, f"step {s} list {LISTS[l]}", ctol)
            row[LISTS[l]] = len(ca["master"])
        counts.append(row)
        e = compare_state(gpu.get_particles(), ref.get_particles(), ftol)
        for k, v in e.items():
            worst[k] = max(worst.get(k, 0.0), v)
        if log:
            log(f"step {s}: contacts {row} errs " + " ".join(f"{k}={v:.1e}" for k, v in e.items()))
    return counts, worst
