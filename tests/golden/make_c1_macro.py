"""Writes tests/golden/c1_macro.json: the macroscopic course of the 1,000-particle settling run
(superellipsoids 5/4/3 mm, exponent 4, 40^3 SDF at 0.5 mm, 500 nodes, closed 0.1 x 0.1 x 0.3 m box,
kn 1e5, kt 0.4 kn, e 0.5, mu 0.5, dt 1e-5, 2,000 steps) computed by the oracle (the serial C++
restatement of the reference kernels; OpenMP over master particles, which leaves its results
unchanged).  Start: a jittered 11 mm lattice with uniform random orientations; orientations of
particles that overlap a neighbour or a wall by more than 5 um are resampled (overlaps detected by
the oracle itself, so the GPU test rebuilds the same start without a GPU-side dependency).  The
particles start agitated (velocities N(0, 0.5 m/s) per axis, angular velocities N(0, 50 rad/s),
seeded): the plain drop makes no contact before about step 1,500, while an agitated start makes a
granular gas that cools by inelastic, frictional collisions within the 2,000 steps, so kinetic-energy
decay, coordination and packing are exercised while individual trajectories diverge.

Run from the repository root:  python tests/golden/make_c1_macro.py [threads]
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import __graft_entry__  # noqa: E402

__graft_entry__.load_package()
import lsdem_amd.scenes as S  # noqa: E402
from oracle_binding import oracle_sim  # noqa: E402

STEPS = 2000
EVERY = 100
LATTICE = 11e-3


def c1_start():
    """The C1 start used by the fixture and by the GPU test."""
    import numpy as np
    sc = S.scene_c1(n=1000, lattice=LATTICE)
    S.remove_initial_overlaps(lambda c: oracle_sim(c), sc)
    rng = np.random.default_rng(20261207)
    sc.particles["vel"] = rng.normal(0.0, 0.5, (sc.n_particles, 3))
    sc.particles["angvel"] = rng.normal(0.0, 50.0, (sc.n_particles, 3))
    return sc


def record(sim, sc, step):
    st = S.macro_stats(sc, sim.get_particles(), [sim.contacts(l) for l in range(5)])
    return {"step": step, "kinetic_energy": st["kinetic_energy"], "solid_fraction": st["solid_fraction"],
            "coordination": st["coordination"], "wall_contacts_per_particle": st["wall_contacts_per_particle"],
            "pp_node_contacts": int(len(sim.contacts(0)["master"]))}


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    t0 = time.perf_counter()
    sc = c1_start()
    sim = sc.load(oracle_sim(sc.cfg, threads=threads))
    rows = [record(sim, sc, 0)]
    for k in range(EVERY, STEPS + 1, EVERY):
        sim.step(EVERY)
        rows.append(record(sim, sc, k))
        print(json.dumps(rows[-1]), flush=True)
    out = {"workload": "C1: 1,000 superellipsoids 5/4/3 mm exponent 4, SDF 40^3 at 0.5 mm, 500 nodes, box "
                       "0.1 x 0.1 x 0.3 m, dt 1e-5, 2,000 steps", "lattice": LATTICE, "seed": sc.meta["seed"],
           "every": EVERY, "steps": STEPS, "agitation": "vel N(0, 0.5 m/s), angvel N(0, 50 rad/s), seed 20261207", "oracle_seconds": time.perf_counter() - t0, "threads": threads,
           "rows": rows}
    (HERE / "c1_macro.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
