"""Benchmark: particle-steps/s of the LS-DEM timestep at 200k particles on one MI355X, plain path
and surrogate-culled path on the same GPU.

Workload (200k particles, config 2 of the benchmark set): 200,000 superquadrics of 10 types
(exponents U[2,8], axis ratios (1, U[.5,1], U[.5,1]), equivalent radius lognormal(median 3 mm,
sigma 0.25)), FP32 SDF grids of 50^3 and 512 nodes per type, periodic x/y on a 0.3 m cell with a
floor wall, linear spring-dashpot (kn 1e5, kt 0.4 kn, e 0.5, mu 0.5), dt 5e-6, seeded
random-lattice start (layer pitch 0.88 x the layer's largest bounding diameter, orientations
resampled until no two particles overlap).  W warm-up steps, then K timed steps; the host clock
brackets the timed steps (lsdem_step returns after the stream is drained).  Per-phase GPU times
come from HIP events recorded on the library's stream inside the timed region, on every 10th step
(each event record costs GPU time of its own; timing every step would add ~13 us to each).
`graph_replay`: the same workload in a second simulation from the same start, W eager warm-up steps,
then the timed steps replayed from a captured 100-step HIP graph (one launch per 100 steps, no
events inside; the capture's own replay untimed): a second throughput figure beside `value`.

Surrogate path: a second context from the same initial state trains the culling network on the
GPU (not timed in the step rate; reported), then runs W + K steps with the network culling node
clusters and an exact audit every 100 steps (audits are inside the timed steps).

Roofline.  Algorithmic bytes per step, 4 B per FP32/int32 scalar, with P candidate pairs, Q the
boundary nodes of the masters of the evaluated pairs (what a node-by-node scan of those pairs
reads), S exact SDF queries, C contacts and N particles, all counted by the kernels:
    k_pair_check     P x 18 x 4                         (pair record, both poses, map)
    k_candidates     Q x 3 x 4 + S x 8 x 4 + C x 16 x 4  (nodes, 8 corners, contact + history)
                     + C x 45 x 4                        (contact law: record, both bodies' state)
    k_reduce_drift   N x (78 + 6) x 4                    (kick + drift + kick, force/torque sums)
achieved = bytes / that kernel's time (HIP events on the library's stream on every 10th timed
step); peak 8 TB/s.  `roofline` is the kernel with the largest time; `kernels` lists all three and
`narrow_phase` the pair check and candidate kernel together.
CPU baseline: the oracle (C++, FP64, pair broad phase with identical contact sets) on the GPU
run's own state at the end of the timed steps (all 200k particles), one untimed step for the
pre-loop build and history, then K steps with OpenMP over all host cores (the contract's value)
and one step on one core.
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

B_PAIR = 18 * 4
B_NODE = 3 * 4
B_QUERY = 8 * 4
B_CONTACT = 16 * 4
B_LAW = 45 * 4
B_BODY = (78 + 6) * 4
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch at the 200k workload, from rocprofv3 FETCH_SIZE + WRITE_SIZE passes scaled by a
# plain copy (profiles/c2_200k_pmc_traffic.md); None for a kernel not measured at its current code.
KERNEL_TRAFFIC_BYTES = {"k_pair_check": 58.61e6, "k_candidates": 12.79e6, "k_reduce_drift": 73.34e6}
METRIC = "particle-steps/s at 200k LS-DEM particles: plain vs surrogate-culled, same GPU"
PROFILE_EVERY = 10


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_bytes(prof: dict, n: int, K: int) -> dict:
    """Algorithmic bytes per step of the three per-step kernels (counts averaged over K steps)."""
    P, Q = prof["pairs"] / K, prof["master_nodes"] / K
    S, C = prof["sdf_queries"] / K, prof["contacts"] / K
    return {"k_pair_check": B_PAIR * P,
            "k_candidates": B_NODE * Q + B_QUERY * S + B_CONTACT * C + B_LAW * C,
            "k_reduce_drift": B_BODY * n}


def rooflines(prof: dict, n: int) -> dict:
    """Per-kernel rooflines from the sampled phase times (pair check: phase pair_check; node culls,
    exact test and contact law: candidates; body pass: reduce)."""
    K = max(prof["steps"], 1)
    KS = max(prof["sampled_steps"], 1)
    ms = {"k_pair_check": prof["ms"]["pair_check"] / KS, "k_candidates": prof["ms"]["candidates"] / KS,
          "k_reduce_drift": prof["ms"]["reduce"] / KS}
    by = kernel_bytes(prof, n, K)
    out = {}
    for k in ms:
        a = by[k] / (ms[k] * 1e-3) / 1e9 if ms[k] > 0 else 0.0
        out[k] = {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
                  "bytes_per_launch": by[k], "ms_per_launch": ms[k], "traffic": KERNEL_TRAFFIC_BYTES[k]}
    nb = by["k_pair_check"] + by["k_candidates"]
    nms = ms["k_pair_check"] + ms["k_candidates"]
    na = nb / (nms * 1e-3) / 1e9 if nms > 0 else 0.0
    tr = [KERNEL_TRAFFIC_BYTES[k] for k in ("k_pair_check", "k_candidates")]
    narrow = {"bound": "hbm", "achieved": na, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": na / HBM_PEAK_GBS,
              "bytes_per_step": nb, "ms_per_step": nms, "traffic": None if None in tr else sum(tr)}
    return {"kernels": out, "narrow_phase": narrow}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def cpu_baseline(sc, state: dict, omp_steps: int, serial_steps: int):
    """The oracle on the GPU run's state (all particles): one untimed step (pre-loop build and the
    history it starts), then omp_steps steps on all host cores and serial_steps on one."""
    import ctypes as C
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_binding import oracle_sim, oracle_lib  # the checker, timed here as the CPU baseline
    cores = host_cores()
    n = sc.n_particles
    sim = sc.load(oracle_sim(sc.cfg, threads=cores))
    sim.set_particles(**state)
    t = time.perf_counter()
    sim.step(1)
    first = time.perf_counter() - t
    t = time.perf_counter()
    sim.step(omp_steps)
    t_omp = time.perf_counter() - t
    serial = None
    if serial_steps > 0:
        oracle_lib().lsdem_oracle_set_threads(sim.ctx, C.c_int(1))
        t = time.perf_counter()
        sim.step(serial_steps)
        t_ser = time.perf_counter() - t
        serial = {"value": n * serial_steps / t_ser, "unit": "particle-steps/s", "cores": 1, "steps": serial_steps,
                  "seconds": t_ser}
    sim.close()
    return {"value": n * omp_steps / t_omp, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"oracle (C++, FP64, pair broad phase: the reference's contact sets), OpenMP over master "
                      f"particles on {cores} threads, the GPU run's state after its timed steps ({n} particles), "
                      f"{omp_steps} steps in {t_omp:.1f} s after 1 untimed step ({first:.1f} s)",
            "serial": serial, "cpu": cpu_model(), "nproc": os.cpu_count()}


def run(L, sc, caps, warmup, steps, label, surrogate=None):
    sim = sc.load(L.Sim({**sc.cfg, **caps}))
    extra = {}
    if surrogate is not None:
        t = time.perf_counter()
        rep = sim.surrogate_train(**surrogate)
        extra["train"] = rep
        log(f"[bench] {label}: trained in {time.perf_counter() - t:.1f}s (GPU {rep['train_ms']:.0f} ms): "
            f"loss {rep['final_loss']:.2e} max err {rep['max_error']:.3e} band {rep['band']:.3e}")
        sim.surrogate_enable(True)
    done = 0
    t = time.perf_counter()
    while done < warmup:
        k = min(1000, warmup - done)
        sim.step(k)
        done += k
        c = sim.counters()
        log(f"[bench] {label} warmup {done}/{warmup} ({time.perf_counter() - t:.1f}s) pairs={c['pairs']} "
            f"contacts={c['contacts']} node_tests={c['node_tests']}")
    # phases timed with HIP events on every PROFILE_EVERY-th step of the timed region (an event
    # record costs GPU time of its own); work counters cover every step
    sim.profile_begin((steps + PROFILE_EVERY - 1) // PROFILE_EVERY, every=PROFILE_EVERY)
    t0 = time.perf_counter()
    sim.step(steps)
    t1 = time.perf_counter()
    prof = sim.profile_end()
    if surrogate is not None:
        extra["report"] = sim.surrogate_report()
    return sim, t1 - t0, prof, extra


def drums(args):
    """Independent rotating drums (particle count lognormal, median 4,000, capped by what the drum
    holds; superellipsoids with exponent U[2,6] and semi-axes 3-5 mm, SDF 32^3, 300 nodes; drum radius
    0.1 m, depth 0.05 m, 1-30 rpm; mu U[0.1,0.8], e U[0.3,0.9]; dt 1e-5) stepped concurrently from one
    host thread each (each context has its own stream), then one after another."""
    import threading
    import __graft_entry__ as g
    g.load_package()
    import lsdem_amd.lsdem as L
    import lsdem_amd.scenes as S
    legs = set(args.drum_legs.split(","))
    rng = np.random.default_rng(args.seed)
    scs = []
    for k in range(args.drums):
        n = int(np.clip(rng.lognormal(np.log(4000.0), 0.8), 500, 20000))
        scs.append(S.scene_drum(n=n, seed=args.seed + k, rpm=float(rng.uniform(1, 30)), mu=float(rng.uniform(0.1, 0.8)),
                                restitution=float(rng.uniform(0.3, 0.9))))
    counts = [sc.n_particles for sc in scs]
    log(f"[bench] drums: {args.drums}, particles {counts}")
    # separate broad-phase launches (concurrent contexts never wait on each other's grid barriers), or
    # the single-launch broad phase on a grid small enough that every context's blocks fit at once
    bp = ({"broad_phase_launches": 1} if args.drum_broad_blocks < 0 else
          {"broad_phase_launches": 0, "broad_phase_blocks": args.drum_broad_blocks})
    make = lambda sc: sc.load(L.Sim({**sc.cfg, **bp}))
    sims = [make(sc) for sc in scs]
    for sim in sims:
        sim.step(10)  # first steps (builds) outside the timing
    errs = []

    def run(sim):
        try:
            sim.step(args.drum_steps)
        except Exception as ex:
            errs.append(ex)
    th = [threading.Thread(target=run, args=(sim,)) for sim in sims]
    t = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    conc = time.perf_counter() - t
    if errs:
        raise errs[0]
    del sims
    conc_graph, chunk = None, 100
    if "graph" in legs:
        # the same, each drum's steps replayed from a captured HIP graph of `chunk` steps
        sims = [make(sc) for sc in scs]
        for sim in sims:
            sim.step(10)
            sim.step_graph(chunk)  # capture (and first replay) outside the timing

        def run_graph(sim):
            try:
                for _ in range(args.drum_steps // chunk):
                    sim.step_graph(chunk)
            except Exception as ex:
                errs.append(ex)
        th = [threading.Thread(target=run_graph, args=(sim,)) for sim in sims]
        t = time.perf_counter()
        for x in th:
            x.start()
        for x in th:
            x.join()
        conc_graph = time.perf_counter() - t
        if errs:
            raise errs[0]
        del sims
    seq = None
    if "sequential" in legs:
        sims = [make(sc) for sc in scs]
        for sim in sims:
            sim.step(10)
        t = time.perf_counter()
        for sim in sims:
            sim.step(args.drum_steps)
        seq = time.perf_counter() - t
        del sims
    total = sum(counts) * args.drum_steps
    tuned = {}
    if args.autotune:
        # one setting per size bucket, measured on the bucket's largest drum stepped alone, then the
        # concurrent run again with each drum under its bucket's setting
        import lsdem_amd.autotune as A
        base = bp
        t = time.perf_counter()
        chosen, report = A.tune_buckets(L, scs, edges=tuple(args.tune_edges), base=base)
        tune_s = time.perf_counter() - t
        log(f"[bench] autotune: {tune_s:.1f} s, " + "; ".join(
            f"bucket {b}: drum {r['representative']} ({r['particles']} particles) -> {r['best']}" for b, r in report.items()))
        sims = [sc.load(L.Sim(A.apply(sc, base, c))) for sc, c in zip(scs, chosen)]
        for sim in sims:
            sim.step(10)
        th = [threading.Thread(target=run, args=(sim,)) for sim in sims]
        t = time.perf_counter()
        for x in th:
            x.start()
        for x in th:
            x.join()
        conc_tuned = time.perf_counter() - t
        if errs:
            raise errs[0]
        del sims
        tuned = {"concurrent_tuned": total / conc_tuned, "tuning_s": tune_s,
                 "tuned": {str(b): r["best"] for b, r in report.items()}}
    print(json.dumps({"metric": "aggregate particle-steps/s of independent rotating drums, one GPU",
                      "drums": args.drums, "particles": counts, "steps": args.drum_steps,
                      "concurrent": total / conc, "sequential": total / seq if seq else None,
                      "speedup": seq / conc if seq else None,
                      "concurrent_graph": (sum(counts) * (args.drum_steps // chunk) * chunk) / conc_graph if conc_graph else None,
                      "broad_phase": bp, **tuned, "unit": "particle-steps/s", "dtype": "f64", "data": "synthetic"}))


def ensemble(args):
    """K independent rotating drums in ONE simulation (drums far apart in one domain, one wall body
    each turning at its own speed; per-drum friction and restitution on its particles and wall), so
    one step of 4 launches, replayed from a captured HIP graph, advances every drum.  Drum k's count
    is lognormal (median 4,000, within 500-20,000), scaled so that 20,000 fills half of the drum's
    lattice sites (a 0.1 m x 0.05 m drum holds a few hundred 3-5 mm particles); superellipsoids with
    exponent U[2,6] and semi-axes 3-5 mm, SDF 32^3, 300 nodes; 1-30 rpm; mu U[0.1,0.8], e U[0.3,0.9];
    dt 1e-5.  The autotuner first picks block size, narrow batch and skin (hence the search-cell
    edge) for the ensemble's simulation by timing each setting on it (--ensemble-no-tune: defaults).
    Against it: the same drums as separate simulations with the defaults, one after another."""
    import __graft_entry__ as g
    g.load_package()
    import lsdem_amd.lsdem as L
    import lsdem_amd.scenes as S
    K = args.ensemble
    rng = np.random.default_rng(args.seed)
    counts = np.clip(rng.lognormal(np.log(4000.0), 0.8, K), 500, 20000)
    sc = S.scene_drum_ensemble(counts, rng.uniform(1, 30, K), rng.uniform(0.1, 0.8, K), rng.uniform(0.3, 0.9, K),
                               seed=args.seed, max_count=20000)
    sizes = np.bincount(sc.meta["drum_of"], minlength=K)
    n = sc.n_particles
    log(f"[bench] ensemble: {K} drums, {n} particles (per drum {sizes.min()}-{sizes.max()}, median "
        f"{int(np.median(sizes))})")
    cfg = dict(sc.cfg)
    for kv in args.ensemble_cfg:  # lsdem_config overrides, e.g. broad_phase_launches=1
        k, v = kv.split("=", 1)
        cfg[k] = float(v) if "." in v or "e" in v else int(v)
    tuned = {}
    if not args.ensemble_no_tune:
        # the configuration's autotuner: one setting for the ensemble's simulation, measured on it
        import lsdem_amd.autotune as A
        t = time.perf_counter()
        base = {k: v for k, v in cfg.items() if k not in sc.cfg}
        best, times = A.tune(L, sc, base=base)
        tuned = {"tuned": best, "tuning_s": time.perf_counter() - t}
        cfg = A.apply(sc, base, best)
        log(f"[bench] autotune: {tuned['tuning_s']:.1f} s -> {best}")
    chunk = args.ensemble_chunk  # 0: eager steps only
    steps = (args.drum_steps // chunk) * chunk if chunk else args.drum_steps
    sim = sc.load(L.Sim(cfg))
    sim.step(10)  # first steps (builds) outside the timing
    graph_s = None
    if chunk:
        sim.step_graph(chunk)  # capture and first replay outside the timing
        t = time.perf_counter()
        for _ in range(steps // chunk):
            sim.step_graph(chunk)
        graph_s = time.perf_counter() - t
        log(f"[bench] ensemble graph replays: {1e3 * graph_s / steps:.4f} ms/step")
    t = time.perf_counter()
    sim.step(steps)
    eager_s = time.perf_counter() - t
    # work and phase split of a further stretch of eager steps (events on every 10th step)
    m = max(10, steps // 4)
    sim.profile_begin((m + 9) // 10, every=10)
    sim.step(m)
    prof = sim.profile_end()
    work = {k: prof[k] / m for k in ("pairs", "evaluated_pairs", "master_nodes", "cluster_tests", "node_tests",
                                     "sdf_queries", "contacts", "rebuilds")}
    phases = {k: v / max(1, prof["sampled_steps"]) for k, v in prof["ms"].items()}
    lists = {name: len(sim.contacts(l)["master"]) for l, name in ((0, "pp"), (1, "pw"))}
    log(f"[bench] ensemble work per step: {work}; node contacts {lists}")
    del sim
    seq = None
    if args.ensemble_seq_steps > 0:
        # the same drums one after another, each a simulation of its own on the plain path
        m = args.ensemble_seq_steps
        t_all = 0.0
        for k in range(K):
            one = S.scene_drum_single(sc, k)
            s1 = one.load(L.Sim(one.cfg))
            s1.step(10)
            t = time.perf_counter()
            s1.step(m)
            t_all += time.perf_counter() - t
            del s1
        seq = n * m / t_all
        log(f"[bench] sequential: {K} drums x {m} steps in {t_all:.2f} s")
    timed_s = graph_s if graph_s is not None else eager_s
    value = n * steps / timed_s
    print(json.dumps({"metric": "aggregate particle-steps/s of independent rotating drums in one simulation, one GPU",
                      "value": value, "unit": "particle-steps/s", "drums": K, "particles": int(n),
                      "per_drum": {"min": int(sizes.min()), "median": float(np.median(sizes)), "max": int(sizes.max())},
                      "steps": steps, "ms_per_step": 1e3 * timed_s / steps, "graph_chunk": chunk,
                      "eager": n * steps / eager_s, "sequential": seq,
                      "sequential_sample_steps": args.ensemble_seq_steps,
                      "speedup_vs_sequential": value / seq if seq else None, **tuned,
                      "work_per_step": work, "phases_ms_per_step": phases, "node_contacts": lists,
                      "dtype": "f64", "data": "synthetic"}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=5000)
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--cpu-steps", type=int, default=30, help="oracle steps timed on all host cores")
    ap.add_argument("--cpu-serial-steps", type=int, default=4, help="oracle steps timed on one core")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--sur-iters", type=int, default=3000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB ten times (HBM byte-counter calibration)")
    ap.add_argument("--gap", type=float, default=-0.12,
                    help="lattice pitch relative to the layer's largest bounding diameter, minus 1 "
                         "(negative: denser than bounding spheres; overlapping orientations are resampled)")
    ap.add_argument("--drums", type=int, default=0, metavar="K",
                    help="instead of the main line: K independent rotating drums stepped concurrently from K host "
                         "threads, then one after another; prints their aggregate particle-steps/s")
    ap.add_argument("--drum-steps", type=int, default=2000)
    ap.add_argument("--drum-broad-blocks", type=int, default=0,
                    help="with --drums: each drum's broad phase as one launch of B blocks (0: the library's choice, "
                         "8 blocks at these sizes; grid barriers, so all drums' blocks must fit on the GPU "
                         "together), or as separate launches (-1)")
    ap.add_argument("--drum-legs", default="concurrent,sequential",
                    help="with --drums: which runs besides the concurrent one to make (graph: concurrent replays of "
                         "captured HIP graphs, one host thread per drum, not run by default: a host-side fault was "
                         "once seen in that mode under a profiler's kernel trace and its cause is not established; "
                         "sequential: one drum after another)")
    ap.add_argument("--no-graph-replay", dest="graph_replay", action="store_false",
                    help="skip the second run of the workload replayed from captured 100-step graphs")
    ap.add_argument("--cfg", nargs="*", default=[], metavar="KEY=VALUE",
                    help="lsdem_config entries of the 200k run (tuning knobs such as narrow_blocks)")
    ap.add_argument("--ensemble", type=int, default=0, metavar="K",
                    help="instead of the main line: K independent rotating drums in one simulation, stepped by "
                         "graph replays (--drum-steps steps), against the same drums one after another")
    ap.add_argument("--ensemble-chunk", type=int, default=100,
                    help="with --ensemble: steps per captured graph (0: eager steps only)")
    ap.add_argument("--ensemble-cfg", nargs="*", default=[], metavar="KEY=VALUE",
                    help="with --ensemble: lsdem_config entries of the ensemble's simulation")
    ap.add_argument("--ensemble-no-tune", action="store_true",
                    help="with --ensemble: run with the default settings instead of the tuned ones")
    ap.add_argument("--ensemble-seq-steps", type=int, default=500,
                    help="with --ensemble: steps per drum of the sequential comparison (0: skip it)")
    ap.add_argument("--autotune", action="store_true",
                    help="with --drums: tune block size, narrow batch and skin per size bucket, then run again")
    ap.add_argument("--tune-edges", type=int, nargs="*", default=[450],
                    help="particle counts that split the drums into size buckets for --autotune")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.empty(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.calibrate_copy} MiB x 10")
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.ensemble:
        ensemble(args)
        return
    if args.drums:
        drums(args)
        return

    import __graft_entry__ as g
    g.load_package()
    import lsdem_amd.lsdem as L
    import lsdem_amd.scenes as S

    t = time.perf_counter()
    sc = S.scene_c2(n=args.n, seed=args.seed, gap=args.gap)
    log(f"[bench] scene C2 n={args.n} built in {time.perf_counter() - t:.1f}s")
    caps = {"pair_capacity": 48 * args.n + 65536, "contact_capacity": 48 * args.n + 65536}
    for kv in args.cfg:  # lsdem_config tuning entries, e.g. narrow_blocks=512
        k, v = kv.split("=", 1)
        caps[k] = float(v) if "." in v or "e" in v else int(v)
    if args.gap < 0:
        t = time.perf_counter()
        moved = S.remove_initial_overlaps(lambda c: L.Sim({**c, **caps}), sc, log=log)
        log(f"[bench] initial overlaps removed ({moved} relocated, {time.perf_counter() - t:.1f}s); "
            f"domain {tuple(float(v) for v in sc.cfg['domain_max'])}")

    sim, wall, prof, _ = run(L, sc, caps, args.warmup, args.steps, "plain")
    ms_step = 1e3 * wall / args.steps
    value = args.n * args.steps / wall
    K = max(prof["steps"], 1)
    KS = max(prof["sampled_steps"], 1)  # steps whose phases were timed
    roof = rooflines(prof, args.n)
    gpu_ms = sum(prof["ms"].values()) / KS
    log(f"[bench] plain: {args.steps} steps in {wall:.2f}s ({ms_step:.4f} ms/step, GPU phases {gpu_ms:.4f} ms/step)")
    log("[bench] phases ms/step: " + " ".join(f"{k}={v / KS:.4f}" for k, v in prof["ms"].items()))
    log(f"[bench] per step: pairs={prof['pairs'] / K:.0f} master_nodes={prof['master_nodes'] / K:.0f} "
        f"cluster_tests={prof['cluster_tests'] / K:.0f} node_tests={prof['node_tests'] / K:.0f} "
        f"sdf_queries={prof['sdf_queries'] / K:.0f} contacts={prof['contacts'] / K:.0f} "
        f"evaluated_pairs={prof['evaluated_pairs'] / K:.0f} rebuilds={prof['rebuilds']}")
    log("[bench] rooflines: " + " ".join(f"{k}={v['frac']:.3f}" for k, v in roof["kernels"].items()) +
        f" narrow={roof['narrow_phase']['frac']:.3f}")
    state = {k: v for k, v in sim.get_particles().items()
             if k in ("pos", "quat", "vel", "angvel", "force", "torque")}
    sim.close()
    del sim
    # the same workload in a second simulation from the same start: W eager warm-up steps, then the K
    # steps replayed from a captured 100-step HIP graph (one launch per 100 steps, no phase events
    # inside): a second throughput figure beside `value`
    graph = None
    if args.graph_replay and args.steps >= 200:
        gsim = sc.load(L.Sim({**sc.cfg, **caps}))
        gsim.step(args.warmup)
        chunk = 100
        n_g = max(chunk, (args.steps // chunk) * chunk)
        gsim.step_graph(chunk)  # capture (its first replay is one of the n_g steps)
        t = time.perf_counter()
        for _ in range(n_g // chunk - 1):
            gsim.step_graph(chunk)
        gs = time.perf_counter() - t
        timed = n_g - chunk
        graph = {"value": args.n * timed / gs, "ms_per_step": 1e3 * gs / timed, "steps": timed, "chunk": chunk}
        log(f"[bench] graph replays: {timed} steps, {graph['ms_per_step']:.4f} ms/step")
        gsim.close()
        del gsim

    sur = None
    if not args.no_surrogate:
        ssim, swall, sprof, extra = run(L, sc, caps, args.warmup, args.steps, "surrogate",
                                        surrogate={"iterations": args.sur_iters})
        sK = max(sprof["steps"], 1)
        sKS = max(sprof["sampled_steps"], 1)
        rep = extra["report"]
        sur = {"value": args.n * args.steps / swall, "ms_per_step": 1e3 * swall / args.steps,
               "speedup_vs_plain": (args.n * args.steps / swall) / value,
               "narrow_ms_per_step": (sprof["ms"]["pair_check"] + sprof["ms"]["candidates"]) / sKS,
               "cluster_tests_per_step": sprof["cluster_tests"] / sK,
               "evaluated_pairs_per_step": sprof["evaluated_pairs"] / sK,
               "plain_evaluated_pairs_per_step": prof["evaluated_pairs"] / K,
               "plain_cluster_tests_per_step": prof["cluster_tests"] / K,
               "train_ms": rep["train_ms"], "train_iterations": rep["iterations"], "final_loss": rep["final_loss"],
               "max_error": rep["max_error"], "band": rep["band"], "length_scale": rep["length_scale"],
               "audits": rep["audits"], "audited_contacts": rep["audited_contacts"],
               "missed_contacts": rep["missed_contacts"]}
        log(f"[bench] surrogate: {sur['ms_per_step']:.3f} ms/step ({sur['speedup_vs_plain']:.2f}x plain), "
            f"audits {rep['audits']} contacts {rep['audited_contacts']} missed {rep['missed_contacts']}")
        log("[bench] surrogate phases ms/step: " + " ".join(f"{k}={v / sKS:.4f}" for k, v in sprof["ms"].items()))
        ssim.close()
        del ssim

    base = None
    if not args.no_cpu:
        t = time.perf_counter()
        base = cpu_baseline(sc, state, args.cpu_steps, args.cpu_serial_steps)
        log(f"[bench] cpu baseline {base['value']:.1f} particle-steps/s on {base['cores']} cores, serial "
            f"{base['serial']['value'] if base['serial'] else float('nan'):.1f} ({time.perf_counter() - t:.1f}s)")
    top = max(roof["kernels"], key=lambda k: roof["kernels"][k]["ms_per_launch"])
    out = {
        "metric": METRIC,
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
        "data": "synthetic (seeded random-lattice start, 200k superquadrics of 10 types)",
        "config": {"workload": "200k superquadrics (10 types, SDF 50^3 FP32, 512 nodes), periodic x/y 0.3 m "
                               "+ floor, dt 5e-6; value = plain path", "particles": args.n, "seed": args.seed,
                   "lattice_gap": args.gap},
        "roofline": {"kernel": top, **roof["kernels"][top]},
        "kernels": roof["kernels"],
        "narrow_phase": roof["narrow_phase"],
        "cpu_baseline": base,
        "surrogate": sur,
        "phases_ms_per_step": {k: v / KS for k, v in prof["ms"].items()},
        "phases_sampled_every": PROFILE_EVERY,
        "work_per_step": {k: prof[k] / K for k in ("pairs", "evaluated_pairs", "master_nodes", "cluster_tests",
                                                    "node_tests", "sdf_queries", "contacts", "rebuilds")},
        "node_sdf_queries_per_s": prof["sdf_queries"] / K / (ms_step * 1e-3),
        "graph_replay": graph,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
