"""Benchmark of the range-solver DCFR hot path on one MI355X.

A step is one batched solve of the river-subgame workload: 4096 seeded river
subgames (random board, Dirichlet(0.5) ranges, pot log-uniform [20, 400],
stack-to-pot U[0.5, 10], 1-3 bet sizes plus all-in, raise cap 1-3), each
solved with the reference solve() schedule: exploitability checked before the
first iteration and every 5 iterations, stop at 0.3% of the pot or after 300
iterations (--mode fixed: exactly 300 iterations then one exploitability
pass).  Inputs are resident in HBM before the timed region; the timed region
is the device solve only.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "river-subgame solves/sec at <=0.3% pot exploitability; turn DCFR iters/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def algorithmic_bytes(storage, iters, checks):
    """20 B per edge-hand per DCFR iteration (regret read as the opponent's
    strategy, read + written at the traverser's update; strategy sum read +
    written) and 4 B per edge-hand per exploitability pass (strategy sums
    read once as the opponent's average strategy)."""
    return float(np.sum(storage * (20.0 * iters + 4.0 * checks)))


def num_checks(iters, max_iters, interval):
    """Exploitability passes of solve(): the initial one, then every
    `interval` iterations and at max_iterations (solver.rs:123,157)."""
    if interval <= 0:
        return np.ones_like(iters)
    c = 1 + iters // interval
    extra = (iters == max_iters) & (iters % interval != 0)
    return c + extra.astype(iters.dtype)


def cpu_baseline(subgames, iters, mode, interval, seconds_target=15.0):
    """The CPU oracle (restatement of the reference range-solver) timed on
    this host: independent subgames on all cores, one subgame per thread at
    a time, as the reference runs river subgames single-threaded."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as O

    threads = max(1, min(os.cpu_count() or 1, 16))
    blobs = [s.build().serialize() for s in subgames]
    targets = [s.target for s in subgames] if mode == "solve" else None
    # calibrate on a few subgames, then size the sample to ~seconds_target
    t0 = time.time()
    O.solve_many(blobs[:4], iters, 1, None if targets is None else targets[:4])
    one = max((time.time() - t0) / 4, 1e-3)
    n = int(min(len(blobs), max(threads, seconds_target * threads / one)))
    secs, expl, its = O.solve_many(blobs[:n], iters, threads, None if targets is None else targets[:n])
    sched = "to 0.3% pot (checked every 5 iterations), cap" if mode == "solve" else "fixed"
    return {"value": round(n / secs, 3), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of the seeded river subgames, solve() {sched} {iters} iterations "
                      f"(mean {float(its.mean()):.1f}), {threads} threads on '{cpu_model()}'"}


def turn_bench(P, W, iters, seed):
    """Turn DCFR iterations/s on the full turn subgame (4 seeded board cards,
    48 rivers, U[0,1] ranges, pot 60, stack 200, 33%/75%/all-in, 2 raises at
    3x): solve_step iterations only, timed over `iters` iterations after one
    warm-up iteration."""
    s = W.turn_uniform01(seed)
    g = s.build()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], iters + 1)
    b.run_iterations(0, 1)
    t = time.perf_counter()
    b.run_iterations(1, iters + 1)
    secs = time.perf_counter() - t
    kms = b.last_kernel_ms()
    edge_hands = float(g.num_storage())
    abytes = 20.0 * edge_hands * iters
    achieved = abytes / (kms * 1e-3) / 1e9
    log(f"turn: {iters} iterations in {secs * 1e3:.1f} ms ({iters / secs:.1f} it/s), {g.num_nodes()} nodes, "
        f"{edge_hands:.0f} edge-hands")
    return {"value": round(iters / secs, 2), "unit": "DCFR iters/s", "iterations": iters, "nodes": g.num_nodes(),
            "edge_hands": edge_hands, "ms_per_iter": round(secs * 1e3 / iters, 4),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 5)}}


def copy_probe(nbytes):
    """One device-to-device copy of `nbytes` (float32 tensors, after a
    warm-up copy), the known-size kernel the FETCH_SIZE / WRITE_SIZE counters
    of the DCFR kernels are scaled by."""
    import torch
    n = nbytes // 4
    x = torch.rand(n, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4}), flush=True)


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--subgames", type=int, default=4096)
    ap.add_argument("--iterations", type=int, default=300)
    ap.add_argument("--mode", choices=["solve", "fixed"], default="solve")
    ap.add_argument("--seed", type=int, default=20260)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--single-sample", type=int, default=64,
                    help="subgames solved one per launch for the single-subgame yardstick (0: skip)")
    ap.add_argument("--turn-iters", type=int, default=200, help="timed turn DCFR iterations (0: skip)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run one plain device copy of this many bytes (calibrates the HBM byte counters)")
    args = ap.parse_args()

    if args.copy_bytes > 0:
        copy_probe(args.copy_bytes)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); see the sharding notes in DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    import poker_solver_rust_amd as P
    from poker_solver_rust_amd import workloads as W

    t0 = time.time()
    subs = W.river_batch(args.subgames, seed=args.seed, iterations=args.iterations)
    games = [s.build() for s in subs]
    storage = np.array([g.num_storage() for g in games], np.float64)
    pots = np.array([s.pot for s in subs], np.float64)
    targets = (0.003 * pots).astype(np.float32)
    log(f"built {len(games)} games in {time.time() - t0:.1f}s; mean edge-hands {storage.mean():.0f}")

    solver = P.GpuBatchSolver(0)
    t0 = time.time()
    solver.prepare_batch(games, args.iterations, targets if args.mode == "solve" else None)
    log(f"prepared in {time.time() - t0:.1f}s; launch {solver.launch_info()}")
    interval = 5 if args.mode == "solve" else 0

    def step():
        solver.reset()
        return solver.solve(args.iterations, interval)

    for _ in range(args.warmup):
        step()
    times, kernel_ms, expl, iters = [], [], None, None
    for _ in range(args.steps):
        solver.reset()
        t = time.perf_counter()
        expl, iters = solver.solve(args.iterations, interval)
        times.append(time.perf_counter() - t)
        kernel_ms.append(solver.last_kernel_ms())
    step_s = float(np.mean(times))
    value = len(games) / step_s
    iters = iters.astype(np.int64)
    checks = num_checks(iters, args.iterations, interval)
    if args.mode == "fixed":
        checks = np.ones_like(iters)
    abytes = algorithmic_bytes(storage, iters, checks)
    kms = float(np.mean(kernel_ms))
    achieved = abytes / (kms * 1e-3) / 1e9
    converged = float(np.mean(expl <= targets))
    log(f"step {step_s * 1e3:.1f} ms, kernel {kms:.1f} ms, {value:.1f} solves/s, mean iterations {iters.mean():.1f}, "
        f"converged {converged:.4f}, max expl/pot {float((expl / pots).max()) * 100:.3f}%")

    traffic = load_traffic("river_batch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic["bytes_per_launch"] if traffic else None}
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"river_batch: {len(games)} seeded river subgames, "
                               f"{'solve() to 0.3% pot, cap' if args.mode == 'solve' else 'fixed'} "
                               f"{args.iterations} DCFR iterations",
                   "subgames": len(games), "iterations": args.iterations, "mode": args.mode, "seed": args.seed,
                   "mean_edge_hands": round(float(storage.mean()), 1),
                   "mean_iterations": round(float(iters.mean()), 2),
                   "converged_frac": round(converged, 5),
                   "max_expl_pot_pct": round(float((expl / pots).max()) * 100, 4)},
        "roofline": roofline,
        "kernel_ms": round(kms, 3),
        "algorithmic_bytes_per_launch": abytes,
    }
    # the plain single-subgame path on the same subgames: one game per batch,
    # one launch each (the batched launch is the product; this is its yardstick)
    n_single = min(args.single_sample, len(games))
    if n_single > 0:
        t = time.perf_counter()
        for i in range(n_single):
            one = P.GpuBatchSolver(0)
            one.prepare_batch([games[i]], args.iterations, targets[i:i + 1] if args.mode == "solve" else None)
            one.solve(args.iterations, interval)
        single_rate = n_single / (time.perf_counter() - t)
        out["single_subgame"] = {"value": round(single_rate, 3), "unit": "solves/s", "sample": n_single,
                                 "batched_over_single": round(value / single_rate, 2)}
        log(f"single-subgame path: {single_rate:.1f} solves/s on {n_single} subgames; batched is "
            f"{value / single_rate:.1f}x")
    if args.turn_iters > 0:
        out["turn"] = turn_bench(P, W, args.turn_iters, args.seed)
    if not args.no_cpu_baseline:
        n_sample = min(len(subs), 512)
        out["cpu_baseline"] = cpu_baseline(subs[:n_sample], args.iterations, args.mode, interval)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
