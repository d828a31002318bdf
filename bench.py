"""Benchmark: particle-steps/s of the LS-DEM timestep at 200k particles (config 2) on one MI355X.

Workload: 200,000 superquadrics of 10 types (exponents U[2,8], axis ratios (1, U[.5,1],
U[.5,1]), equivalent radius lognormal(median 3 mm, sigma 0.25)), FP32 SDF grids of 50^3 and 512
nodes per type, periodic x/y on a 0.3 m cell with a floor wall, linear spring-dashpot
(kn 1e5, kt 0.4 kn, e 0.5, mu 0.5), dt 5e-6.  Synthetic random-lattice start (seeded).
W warm-up steps, then K timed steps; the host clock brackets the timed steps (lsdem_step
returns after the stream is drained).  Per-phase GPU times come from HIP events recorded on
the library's stream inside the timed region.

The roofline entry is for the dominant kernel, the narrow phase (k_narrow).  Its algorithmic
bytes per launch are counted from the step's own work counters:
    per candidate pair  B_PAIR   bytes (pair record, both bodies' pose/velocity/material, sums)
    per node tested     B_NODE   bytes (node record)
    per SDF query       B_QUERY  bytes (8 FP32 corners)
    per contact         B_CONTACT bytes (history probe + contact record write)
The CPU baseline is the oracle (serial C++, 1 core) on a bounded sample of the same workload:
the same generator at fewer particles with the periodic cell shrunk to keep the density.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

B_PAIR = 8 + 2 * 232 + 76
B_NODE = 32
B_QUERY = 32
B_CONTACT = 36 + 104
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def narrow_bytes(prof: dict) -> float:
    return (B_PAIR * prof["pairs"] + B_NODE * prof["node_tests"] + B_QUERY * prof["sdf_queries"] +
            B_CONTACT * prof["contacts"])


def cpu_baseline(S, n_full: int, n_sample: int, steps: int, seed: int):
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_binding import oracle_sim  # the checker, timed here as the CPU baseline
    cell = 0.3 * float(np.sqrt(n_sample / n_full))
    sc = S.scene_c2(n=n_sample, cell_xy=cell, seed=seed)
    sim = sc.load(oracle_sim(sc.cfg))
    sim.step(1)  # first step builds the history; time the following ones
    t0 = time.perf_counter()
    sim.step(steps)
    dt = time.perf_counter() - t0
    return {"value": n_sample * steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle (serial C++, FP64), C2 generator at {n_sample} particles, periodic cell "
                      f"{cell:.4f} m (same density), {steps} steps after 1 untimed step, {dt:.1f} s",
            "cpu": platform.processor() or platform.machine(), "nproc": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--cpu-sample", type=int, default=5_000)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gap", type=float, default=-0.12,
                    help="lattice pitch relative to the layer's largest bounding diameter, minus 1 "
                         "(negative: denser than bounding spheres; overlapping orientations are resampled)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import __graft_entry__ as g
    g.load_package()
    import lsdem_amd.lsdem as L
    import lsdem_amd.scenes as S

    t = time.perf_counter()
    sc = S.scene_c2(n=args.n, seed=args.seed, gap=args.gap)
    log(f"[bench] scene C2 n={args.n} built in {time.perf_counter() - t:.1f}s; domain {sc.cfg['domain_max']}")
    caps = {"pair_capacity": 48 * args.n + 65536, "contact_capacity": 48 * args.n + 65536}
    if args.gap < 0:
        t = time.perf_counter()
        moved = S.remove_initial_overlaps(lambda c: L.Sim({**c, **caps}), sc, log=log)
        log(f"[bench] initial overlaps removed ({moved} relocated, {time.perf_counter() - t:.1f}s); "
            f"domain {sc.cfg['domain_max']}")
    cfg = {**sc.cfg, **caps}
    sim = sc.load(L.Sim(cfg))
    done = 0
    t = time.perf_counter()
    while done < args.warmup:
        k = min(500, args.warmup - done)
        sim.step(k)
        done += k
        c = sim.counters()
        log(f"[bench] warmup {done}/{args.warmup} ({time.perf_counter() - t:.1f}s) pairs={c['pairs']} "
            f"contacts={c['contacts']} node_tests={c['node_tests']}")
    sim.profile_begin(args.steps)
    t0 = time.perf_counter()
    sim.step(args.steps)
    t1 = time.perf_counter()
    prof = sim.profile_end()
    wall = t1 - t0
    ms_step = 1e3 * wall / args.steps
    value = args.n * args.steps / wall
    K = max(prof["steps"], 1)
    narrow_ms = prof["ms"]["narrow"] / K
    per_launch_bytes = narrow_bytes(prof) / K
    achieved = per_launch_bytes / (narrow_ms * 1e-3) / 1e9
    gpu_ms = sum(prof["ms"].values()) / K
    log(f"[bench] timed {args.steps} steps in {wall:.2f}s ({ms_step:.3f} ms/step, GPU phases {gpu_ms:.3f} ms/step)")
    log("[bench] phases ms/step: " + " ".join(f"{k}={v / K:.4f}" for k, v in prof["ms"].items()))
    log(f"[bench] per step: pairs={prof['pairs'] / K:.0f} cluster_tests={prof['cluster_tests'] / K:.0f} "
        f"node_tests={prof['node_tests'] / K:.0f} "
        f"sdf_queries={prof['sdf_queries'] / K:.0f} contacts={prof['contacts'] / K:.0f}")
    base = None
    if not args.no_cpu:
        t = time.perf_counter()
        base = cpu_baseline(S, args.n, args.cpu_sample, args.cpu_steps, args.seed)
        log(f"[bench] cpu baseline {base['value']:.1f} particle-steps/s ({time.perf_counter() - t:.1f}s)")
    out = {
        "metric": "particle-steps/s at 200k LS-DEM particles (plain path)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random-lattice start of config 2)",
        "config": {"workload": "C2: 200k superquadrics, 10 types, SDF 50^3 FP32, 512 nodes, periodic x/y 0.3 m "
                               "+ floor, dt 5e-6", "particles": args.n, "seed": args.seed, "lattice_gap": args.gap},
        "roofline": {"bound": "hbm", "kernel": "k_narrow", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": per_launch_bytes, "ms_per_launch": narrow_ms},
        "cpu_baseline": base,
        "phases_ms_per_step": {k: v / K for k, v in prof["ms"].items()},
        "work_per_step": {k: prof[k] / K for k in ("pairs", "cluster_tests", "node_tests", "sdf_queries",
                                                    "contacts")},
        "node_sdf_queries_per_s": prof["sdf_queries"] / K / (ms_step * 1e-3),
        "surrogate": None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
