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

BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 MFMA
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


def algorithmic_bytes(storage, iters, checks, finalize=True):
    """20 B per edge-hand per DCFR iteration (regret read as the opponent's
    strategy, read + written at the traverser's update; strategy sum read +
    written), 4 B per edge-hand per exploitability pass (strategy sums read
    once as the opponent's average strategy) and 8 B per edge-hand for
    finalize (strategy sums read in both players' cfvalue traversals)."""
    return float(np.sum(storage * (20.0 * iters + 4.0 * checks + (8.0 if finalize else 0.0))))


def num_checks(iters, max_iters, interval):
    """Exploitability passes of solve(): the initial one, then every
    `interval` iterations and at max_iterations (solver.rs:123,157)."""
    if interval <= 0:
        return np.ones_like(iters)
    c = 1 + iters // interval
    extra = (iters == max_iters) & (iters % interval != 0)
    return c + extra.astype(iters.dtype)


def box_threads():
    return max(1, min(os.cpu_count() or 1, 16))


def cpu_baseline(subgames, iters, mode, interval, seconds_target=15.0):
    """The CPU oracle (restatement of the reference range-solver) timed on
    this host: independent subgames on all cores, one subgame per thread at
    a time, as the reference runs river subgames single-threaded; each solve
    ends with finalize (the root cfvalues), as solve() does."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as O

    threads = box_threads()
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
                      f"(mean {float(its.mean()):.1f}) + finalize, {threads} threads on '{cpu_model()}'"}


def cpu_turn_baseline(game, seconds_target=8.0):
    """The CPU oracle on the same turn subgame on all cores (the turn's
    children in parallel, the reference's rayon policy): DCFR iterations/s of
    plain solve_step iterations, and of the solve() schedule (exploitability
    before the first iteration and every 5), each over a bounded number of
    iterations sized to ~seconds_target."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as O

    threads = box_threads()
    o = O.OracleGame(game.serialize())
    o.set_threads(threads)
    t = time.time()
    o.solve_step(0)
    one = max(time.time() - t, 1e-3)
    n = int(max(5, min(500, seconds_target / one)) // 5 * 5)
    t = time.time()
    for k in range(1, n + 1):
        o.solve_step(k)
    step_rate = n / (time.time() - t)
    o2 = O.OracleGame(game.serialize())
    o2.set_threads(threads)
    t = time.time()
    o2.solve(n, 0.0)
    solve_rate = n / (time.time() - t)
    return {"value": round(step_rate, 3), "unit": "DCFR iters/s", "cores": threads, "kind": "port",
            "with_solve_checks": round(solve_rate, 3),
            "sample": f"{n} solve_step iterations of the same turn subgame ({n} more under the solve() schedule "
                      f"for with_solve_checks), {threads} threads on '{cpu_model()}'"}


def fp64_gap_river(P, games, pots, iters, expl64, n_sample):
    """FP64 exploitability gap on a sample of the batch: the device's FP64
    best response of its own f32 solution against the FP64 oracle's solve of
    the same tree with the same number of iterations."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as O

    n = min(n_sample, len(games))
    blobs = [g.serialize() for g in games[:n]]
    secs, eo = O.fp64_exploitability_many(blobs, iters[:n], box_threads())
    gap = (expl64[:n] - eo) / pots[:n]
    return {"sample": n, "max_abs_gap_pot_pct": round(float(np.abs(gap).max()) * 100, 5),
            "mean_gap_pot_pct": round(float(gap.mean()) * 100, 6),
            "device_fp64_max_expl_pot_pct": round(float((expl64 / pots).max()) * 100, 5),
            "oracle_seconds": round(secs, 2)}


def turn_bench(P, W, iters, seed, cpu=True):
    """Turn DCFR iterations/s on the full turn subgame (4 seeded board cards,
    48 rivers, U[0,1] ranges, pot 60, stack 200, 33%/75%/all-in, 2 raises at
    3x): solve_step iterations only, timed over `iters` iterations after one
    warm-up iteration (a fresh solve of `iters` iterations, reset first);
    then the FP64 exploitability gap of that solve against the FP64 oracle's
    solve of the same tree."""
    s = W.turn_uniform01(seed)
    g = s.build()
    b = P.GpuBatchSolver(0)
    t = time.perf_counter()
    b.prepare_batch([g], iters + 1)
    prep = time.perf_counter() - t
    # the same iterations with every phase launch issued from the host, for comparison
    b.set_launch(0, 1, 1)
    b.run_iterations(0, 1)
    b.reset()
    t = time.perf_counter()
    b.run_iterations(0, iters)
    eager = iters / (time.perf_counter() - t)
    shape = b.autotune(10)
    log(f"turn: autotuned launch {shape}; eager launches {eager:.1f} it/s")
    b.run_iterations(0, 1)
    b.reset()
    t = time.perf_counter()
    b.run_iterations(0, iters)
    secs = time.perf_counter() - t
    kms = b.last_kernel_ms()
    edge_hands = float(g.num_storage())
    abytes = 20.0 * edge_hands * iters
    achieved = abytes / (kms * 1e-3) / 1e9
    log(f"turn: {iters} iterations in {secs * 1e3:.1f} ms ({iters / secs:.1f} it/s), {g.num_nodes()} nodes, "
        f"{edge_hands:.0f} edge-hands")
    out = {"value": round(iters / secs, 2), "unit": "DCFR iters/s", "iterations": iters, "nodes": g.num_nodes(),
           "edge_hands": edge_hands, "ms_per_iter": round(secs * 1e3 / iters, 4), "prepare_s": round(prep, 3),
           "launch": shape, "eager_launches_it_per_s": round(eager, 2),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5)}}
    if cpu:
        out["cpu_baseline"] = cpu_turn_baseline(g)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_binding as O

        e64 = float(b.exploitability_f64()[0])
        t = time.time()
        o = O.OracleGame(g.serialize(), 64)
        o.set_threads(box_threads())
        for k in range(iters):
            o.solve_step(k)
        eo = o.exploitability()
        out["fp64_gap"] = {"device_fp64_exploitability": e64, "oracle_fp64_exploitability": eo,
                           "gap_chips": e64 - eo, "gap_pot_pct": (e64 - eo) / s.pot * 100,
                           "oracle_seconds": round(time.time() - t, 2)}
        log(f"turn fp64 gap: device {e64:.6f} vs oracle {eo:.6f} chips (pot {s.pot})")
    return out


def net_solve_bench(args):
    """The net-backed depth-limited turn solve on the full turn subgame (4
    seeded board cards, U[0,1] ranges, pot 60, stack 200, 33%/75%/all-in, two
    raises at 3x).  Steps: a check-down depth-limited solve records the
    boundary reaches the solve visits; river training solves from the batched
    GPU solver (half on those reaches, half on synthetic boundary-like ranges,
    random river cards); value-net training on the GPU; the turn tree cut at
    the river deal solved with the net at its leaves.  Yardsticks on the same
    tree: check-down leaf values (the method without the net), exact river
    solves at the leaves (the fallback) and the exact full turn solve (the
    target).  Errors are root expected values (query.rs:639-666) as a fraction
    of the pot."""
    import poker_solver_rust_amd as P
    from poker_solver_rust_amd import net_solve as N
    from poker_solver_rust_amd import workloads as W

    turn = W.turn_uniform01(args.seed)
    dl_game = P.PostFlopGame.with_config(turn.card_config(), P.ActionTree(N.depth_limited_tree(turn)))
    pots = sorted({dl_game.boundary_pot(b) for b in range(dl_game.num_boundary_nodes())})
    log(f"turn board {turn.board[:4]}, {dl_game.num_boundary_nodes()} river-deal boundaries, pots {pots}")
    iters = args.turn_solve_iters
    chk = N.CheckdownLeaves(turn.board[:4])
    snaps = []
    t0 = time.time()
    g_chk, b_chk, leaf_chk, dcfr_chk = N.solve_depth_limited(turn, chk, iters, 10, log, snapshots=snaps)
    t_chk = time.time() - t0
    situations = [it for items in snaps for it in items]
    t0 = time.time()
    ts = N.river_training_set(turn, pots, args.net_train_solves, seed=args.seed + 1, iterations=args.iterations,
                              log=log, situations=situations, checkdown=chk)
    t_data = time.time() - t0
    # one net per training seed on the same data; the ensemble averages their corrections
    seeds = [args.seed + 97 * k for k in range(max(1, args.net_seeds))]
    nets, per_seed = [], []
    t0 = time.time()
    for sd in seeds:
        net, losses, mae = N.train_value_net(ts, epochs=args.net_epochs, seed=sd)
        nets.append(net)
        per_seed.append({"seed": sd, "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1]),
                         "holdout_mae_half_pots": mae})
        log(f"net seed {sd}: loss {losses[0]:.4f} -> {losses[-1]:.4f}, hold-out MAE {mae:.4f} half-pots")
    t_train = time.time() - t0
    dev = N.DeviceCheckdownLeaves(turn.board[:4], dl_game.num_boundary_nodes())
    for k, net in enumerate(nets):   # each net alone, to show the seed-to-seed spread
        _, b_one, _, _ = N.solve_depth_limited(turn, N.NetLeaves(net, dev), iters, 10, lambda m: None)
        per_seed[k]["b"] = b_one
    t0 = time.time()
    g_net, b_net, leaf_net, dcfr_net = N.solve_depth_limited(turn, N.NetLeaves(nets, dev), iters, 10, log)
    t_net = time.time() - t0
    t0 = time.time()
    g_exl, b_exl, leaf_exl, dcfr_exl = N.solve_depth_limited(turn, N.ExactLeaves(turn, args.iterations), iters, 10,
                                                              log)
    t_exl = time.time() - t0
    g_ex = turn.build()
    b_ex = P.GpuBatchSolver(0)
    b_ex.prepare_batch([g_ex], iters)
    t0 = time.time()
    b_ex.run_iterations(0, iters)
    t_ex = time.time() - t0
    err = {name: N.root_ev_error(b_ex, b, g_ex, turn.pot)
           for name, b in (("net_backed", b_net), ("checkdown", b_chk), ("exact_leaves", b_exl))}
    for ps in per_seed:
        e = N.root_ev_error(b_ex, ps.pop("b"), g_ex, turn.pot)
        ps["game_value_error"] = max(e["p0_value_diff_pot"], e["p1_value_diff_pot"])
        log(f"net seed {ps['seed']} alone: game value error {ps['game_value_error'] * 100:.3f}% of the pot")
    # the net against the same method with exact leaves (isolates the net)
    err["net_vs_exact_leaves"] = N.root_ev_error(b_exl, b_net, g_exl, turn.pot)
    for k, v in err.items():
        log(f"{k}: {v}")
    gv = {k: max(v["p0_value_diff_pot"], v["p1_value_diff_pot"]) for k, v in err.items()}
    # leaf yardstick: leaf values at the exact-leaf solve's final (average
    # strategy) reaches, net and check-down against exact river solves
    items = N.boundary_items(turn, g_exl, b_exl)
    v_ex = N.ExactLeaves(turn, args.iterations).values_for(items)
    leaf = {"net": N.leaf_error(items, N.NetLeaves(nets, dev).values_for(items), v_ex),
            "checkdown": N.leaf_error(items, chk.values_for(items), v_ex)}
    log(f"leaf error (fraction of pot, reach-weighted): {leaf}")
    out = {"metric": "net-backed turn solve: game value error vs the exact turn solve (fraction of pot)",
           "value": gv["net_backed"], "game_value_error": gv,
           "gain_over_checkdown": gv["checkdown"] / max(gv["net_backed"], 1e-12),
           "leaf_error_pot": leaf,
           "unit": "pot fraction", "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
           "dtype": "bf16 net, f32 solve",
           "config": {"workload": "turn_uniform01 seed %d, depth_limit 0 at the river deal" % args.seed,
                      "turn_iterations": iters, "leaf_refresh": 10, "train_solves": ts.solves,
                      "train_solve_iterations_cap": args.iterations, "epochs": args.net_epochs,
                      "net": "2706-512x4(GELU)-2652 bf16", "ensemble_seeds": seeds},
           "per_seed": per_seed,
           **err,
           "training": {"data_seconds": t_data, "solve_seconds": ts.solve_seconds,
                        "river_solves_per_s": ts.solves / max(ts.solve_seconds, 1e-9), "train_seconds": t_train},
           "seconds": {"net_backed_turn": t_net, "net_leaf_eval": leaf_net, "net_dcfr": dcfr_net,
                       "checkdown_turn": t_chk, "exact_leaves_turn": t_exl, "exact_turn": t_ex}}
    print(json.dumps(out), flush=True)


def reduced_precision_bench(args):
    """Reduced-precision turn solve: the seeded turn subgame solved on the device with FP32,
    FP16 and BF16 storage (16-bit words with a per-node FP32 scale, FP32
    arithmetic) for `--half-storage-iters` DCFR iterations; the exploitability of each
    by the device FP64 best response, and the per-hand root cfvalues against
    the FP64 CPU oracle's solve of the identical tree (same iterations)."""
    import poker_solver_rust_amd as P
    from poker_solver_rust_amd import workloads as W

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as O

    s = W.turn_uniform01(args.seed)
    g = s.build()
    iters = args.half_storage_iters
    t = time.time()
    o = O.OracleGame(g.serialize(), 64)
    for k in range(iters):
        o.solve_step(k)
        if (k + 1) % 50 == 0:
            log(f"fp64 oracle: {k + 1}/{iters} iterations")
    ref = [o.root_cfvalues(p) for p in range(2)]
    ref_expl = o.exploitability()
    t_oracle = time.time() - t
    out = {"metric": "reduced-precision turn solve: FP64 exploitability and root cfv error vs the FP64 oracle",
           "config": {"workload": "turn_uniform01 seed %d" % args.seed, "iterations": iters, "nodes": g.num_nodes()},
           "fp64_oracle": {"exploitability": ref_expl, "seconds": t_oracle}, "unit": "chips (pot %d)" % s.pot}
    for storage in ("f32", "fp16", "bf16"):
        b = P.GpuBatchSolver(0, storage=storage)
        b.prepare_batch([g], iters)
        t = time.perf_counter()
        b.run_iterations(0, iters)
        secs = time.perf_counter() - t
        e64 = float(b.exploitability_f64()[0])
        cf = b.root_cfvalues()[0]
        err = max(float(np.abs(cf[p].astype(np.float64) - ref[p]).max()) for p in range(2))
        scale = max(float(np.abs(ref[p]).max()) for p in range(2))
        out[storage] = {"exploitability_fp64": e64, "gap_vs_fp64_oracle": e64 - ref_expl,
                        "root_cfv_max_abs_err": err, "root_cfv_max_abs_err_rel": err / max(scale, 1e-30),
                        "iters_per_s": iters / secs}
        log(f"{storage}: exploitability {e64:.5f} (oracle {ref_expl:.5f}), root cfv max err {err:.3e}, "
            f"{iters / secs:.1f} it/s")
    print(json.dumps(out), flush=True)


def river_uniform_bench(args):
    """One seeded river subgame (uniform ranges, 50%/100%/all-in,
    one raise), 1000 DCFR iterations with the solve() checks every 5 and
    finalize, on the device (single-game path) and by the CPU oracle (one
    thread); the FP64 exploitability gap of the device's solution against the
    FP64 oracle's solve of the same tree."""
    import poker_solver_rust_amd as P
    from poker_solver_rust_amd import workloads as W

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as O

    s = W.river_uniform(args.seed)
    g = s.build()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 1000)
    b.solve(1000, 5)
    b.reset()
    t = time.perf_counter()
    e, it = b.solve(1000, 5)
    secs = time.perf_counter() - t
    o = O.OracleGame(g.serialize())
    t = time.time()
    eo, ito = o.solve(1000, 0.0)   # ends with finalize
    t_cpu = time.time() - t
    e64 = float(b.exploitability_f64()[0])
    o64 = O.OracleGame(g.serialize(), 64)
    for k in range(int(it[0])):
        o64.solve_step(k)
    eo64 = o64.exploitability()
    out = {"metric": "single river subgame, 1000 DCFR iterations: seconds", "config": {
               "workload": "river_uniform seed %d" % args.seed, "nodes": g.num_nodes(),
               "edge_hands": g.num_storage()},
           "device": {"seconds": secs, "exploitability": float(e[0]), "iterations": int(it[0])},
           "cpu_oracle_1_thread": {"seconds": t_cpu, "exploitability": eo, "iterations": ito},
           "fp64_gap": {"device_fp64_exploitability": e64, "oracle_fp64_exploitability": eo64,
                        "gap_chips": e64 - eo64, "gap_pot_pct": (e64 - eo64) / s.pot * 100},
           "higher_is_better": False}
    print(json.dumps(out), flush=True)


def flop_bench(args):
    """Flop-root subgame (three streets on the device: the flop tile, then
    one tile per turn card holding its river deals): DCFR iterations/s after
    one warm-up iteration, the HBM roofline on 20 B per edge-hand per
    iteration, and the CPU oracle (one thread) on the same tree."""
    import poker_solver_rust_amd as P
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as O

    oop, ip = "22+,A2s+,K9s+,QTs+,JTs,ATo+,KJo+", "55+,A8s+,KTs+,QJs,AJo+,KQo"
    cc = P.CardConfig(range=(P.Range.parse(oop), P.Range.parse(ip)), flop=P.flop_from_str("Qs Jh 2c"))
    sizes = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=200, flop_bet_sizes=sizes,
                      turn_bet_sizes=sizes, river_bet_sizes=sizes)
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    iters = args.flop_iters
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], iters + 1)
    shape = b.autotune(3)
    log(f"flop: autotuned launch {shape}")
    b.run_iterations(0, 1)
    t = time.perf_counter()
    b.run_iterations(1, iters + 1)
    secs = time.perf_counter() - t
    kms = b.last_kernel_ms()
    edge_hands = float(g.num_storage())
    achieved = 20.0 * edge_hands * iters / (kms * 1e-3) / 1e9
    o = O.OracleGame(g.serialize())
    o.solve_step(0)
    cpu_iters = 3
    t = time.perf_counter()
    for k in range(1, cpu_iters + 1):
        o.solve_step(k)
    cpu_secs = time.perf_counter() - t
    out = {"metric": "flop-root subgame DCFR iterations/s", "value": round(iters / secs, 2), "unit": "DCFR iters/s",
           "higher_is_better": True, "n_gpus": 1, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"flop Qs Jh 2c, OOP {oop}, IP {ip}, pot 100, stack 200, bets 50%/all-in on "
                                  "every street, no raises", "nodes": g.num_nodes(), "edge_hands": edge_hands,
                      "tiles": g.validate_layout()["tiles"], "iterations": iters, "launch": shape},
           "ms_per_iter": round(secs * 1e3 / iters, 3),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 5)},
           "cpu_baseline": {"value": round(cpu_iters / cpu_secs, 3), "unit": "DCFR iters/s", "cores": 1, "kind": "port",
                            "sample": f"{cpu_iters} solve_step iterations of the same tree after one, 1 thread on "
                                      f"'{cpu_model()}'"}}
    log(f"flop: {iters} iterations in {secs * 1e3:.1f} ms ({iters / secs:.1f} it/s); CPU oracle "
        f"{cpu_iters / cpu_secs:.2f} it/s")
    print(json.dumps(out), flush=True)


def mlp_bench(args):
    """Value-net throughput on the 2706 -> 4 x 512 -> 2652 net (BF16 MFMA):
    inference on 8192 rows, and training epochs at batch 512 and 4096 over a
    32768-row resident set; FLOP counts 2 * sum(d_in * d_out) per row forward
    and 3x that per training row (forward, input and weight gradients)."""
    from poker_solver_rust_amd.value_net import DEFAULT_DIMS, ValueNet

    dims = DEFAULT_DIMS
    fwd_flop = 2.0 * sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
    rng = np.random.default_rng(args.seed)
    n = 32768
    x = rng.normal(0, 1, (n, dims[0])).astype(np.float32)
    y = rng.normal(0, 0.1, (n, dims[-1])).astype(np.float32)
    out = {"metric": "value-net MFMA throughput", "unit": "TFLOP/s", "peak_bf16_dense": BF16_PEAK_TFLOPS,
           "flop_per_row_forward": fwd_flop}
    net = ValueNet(max_batch=8192, seed=1)
    net.forward(x[:8192])
    net.forward(x[:8192])
    ms = net.last_ms()
    # host rows in and out: PCIe and the host-side unpacking are included
    out["forward_8192_rows_from_host"] = {"ms": ms, "tflops": fwd_flop * 8192 / (ms * 1e-3) / 1e12}
    net.set_dataset(x, y, np.ones_like(y))
    for batch in (512, 4096):
        net.train(1, batch, seed=2)
        net.train(2, batch, seed=3)
        ms = net.last_ms()
        rows = 2 * (n // batch) * batch
        tf = 3.0 * fwd_flop * rows / (ms * 1e-3) / 1e12
        out[f"train_batch_{batch}"] = {"ms_per_epoch": ms / 2, "rows_per_s": rows / (ms * 1e-3), "tflops": tf,
                                       "frac_of_peak": tf / BF16_PEAK_TFLOPS}
        log(f"train batch {batch}: {ms / 2:.1f} ms/epoch, {tf:.1f} TFLOP/s")
    log(f"forward 8192 rows from host memory: {out['forward_8192_rows_from_host']}")
    out["cpu_baseline"] = mlp_cpu_baseline(net.get_params(), dims, x[:512], y[:512], fwd_flop)
    log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


def mlp_cpu_baseline(params, dims, x, y, fwd_flop, reps=3):
    """The numpy oracle of the same net (oracle/mlp_oracle.py) on the host in
    FP32: forward, masked Huber loss and every gradient for one batch of
    rows, best of `reps`; numpy's BLAS threads (the box's OMP_NUM_THREADS)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mlp_oracle as M

    p = np.asarray(params, np.float32)
    mask = np.ones_like(y)
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        M.loss_and_grads(p, dims, x, y, mask, 1.0)
        best = min(best, time.perf_counter() - t)
    rows = x.shape[0]
    return {"value": round(3.0 * fwd_flop * rows / best / 1e12, 4), "unit": "TFLOP/s", "rows_per_s": round(rows / best, 1),
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
            "sample": f"one FP32 forward + backward (no optimizer step) of {rows} rows, best of {reps}, on '{cpu_model()}'"}


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
    ap.add_argument("--turn-iters", type=int, default=500, help="timed turn DCFR iterations (0: skip)")
    ap.add_argument("--fp64-sample", type=int, default=128,
                    help="river subgames whose FP64 exploitability gap to the FP64 oracle is measured (0: skip)")
    ap.add_argument("--net-solve", action="store_true",
                    help="run the net-backed depth-limited turn solve instead (river solves -> value net -> turn)")
    ap.add_argument("--net-train-solves", type=int, default=32768, help="river solves in the value-net training set")
    ap.add_argument("--net-epochs", type=int, default=40)
    ap.add_argument("--net-seeds", type=int, default=3, help="value nets trained (ensemble of their corrections)")
    ap.add_argument("--turn-solve-iters", type=int, default=500)
    ap.add_argument("--half-storage", action="store_true",
                    help="reduced-precision turn solve (FP16 / BF16 storage, FP64 oracle) instead")
    ap.add_argument("--half-storage-iters", type=int, default=500)
    ap.add_argument("--mlp", action="store_true", help="value-net MFMA throughput instead")
    ap.add_argument("--flop", action="store_true", help="flop-root subgame DCFR iterations/s instead")
    ap.add_argument("--flop-iters", type=int, default=50)
    ap.add_argument("--river-single", action="store_true",
                    help="single river subgame, 1000 iterations, device and CPU oracle instead")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run one plain device copy of this many bytes (calibrates the HBM byte counters)")
    args = ap.parse_args()

    if args.copy_bytes > 0:
        copy_probe(args.copy_bytes)
        return
    if args.half_storage:
        reduced_precision_bench(args)
        return
    if args.river_single:
        river_uniform_bench(args)
        return
    if args.mlp:
        mlp_bench(args)
        return
    if args.flop:
        flop_bench(args)
        return
    if args.net_solve:
        net_solve_bench(args)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the sharding across GPUs is a design note, not built",
              file=sys.stderr)
        sys.exit(2)

    import poker_solver_rust_amd as P
    from poker_solver_rust_amd import workloads as W

    t0 = time.time()
    subs = W.river_batch(args.subgames, seed=args.seed, iterations=args.iterations)
    games = W.build_all(subs)
    storage = np.array([g.num_storage() for g in games], np.float64)
    pots = np.array([s.pot for s in subs], np.float64)
    targets = (0.003 * pots).astype(np.float32)
    log(f"built {len(games)} games in {time.time() - t0:.1f}s; mean edge-hands {storage.mean():.0f}")

    solver = P.GpuBatchSolver(0)
    t0 = time.time()
    solver.prepare_batch(games, args.iterations, targets if args.mode == "solve" else None)
    prepare_s = time.time() - t0
    shape = solver.autotune(5)
    log(f"prepared in {prepare_s:.1f}s; launch {solver.launch_info()}, autotuned {shape}")
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
        # one launch: solve() schedule, then finalize (root cfvalues)
        expl, iters = solver.solve(args.iterations, interval)
        times.append(time.perf_counter() - t)
        kernel_ms.append(solver.last_kernel_ms())
    step_s = float(np.mean(times))
    value = len(games) / step_s
    iters = iters.astype(np.int64)
    checks = num_checks(iters, args.iterations, interval)
    if args.mode == "fixed":
        checks = np.ones_like(iters) * 2
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
                               f"{args.iterations} DCFR iterations, then finalize",
                   "subgames": len(games), "iterations": args.iterations, "mode": args.mode, "seed": args.seed,
                   "launch": {**solver.launch_info(), **shape},
                   "mean_edge_hands": round(float(storage.mean()), 1),
                   "mean_iterations": round(float(iters.mean()), 2),
                   "converged_frac": round(converged, 5),
                   "max_expl_pot_pct": round(float((expl / pots).max()) * 100, 4)},
        "roofline": roofline,
        "kernel_ms": round(kms, 3),
        "algorithmic_bytes_per_launch": abytes,
        "prepare_s": round(prepare_s, 3),
        "end_to_end_solves_per_s": round(len(games) / (prepare_s + step_s), 3),
    }
    if traffic:
        out["traffic_measured_on"] = traffic.get("measured_on")
    if args.fp64_sample > 0:
        e64 = solver.exploitability_f64()
        out["fp64_gap"] = fp64_gap_river(P, games, pots, iters, e64, args.fp64_sample)
        log(f"fp64 gap: {out['fp64_gap']}")
    if args.mode == "solve":
        # the same batch at exactly 300 iterations (no early stop), then one
        # exploitability pass and finalize
        fixed = 300
        solver.set_targets(None)
        solver.reset()
        t = time.perf_counter()
        fe, fi = solver.solve(fixed, 0)
        fsecs = time.perf_counter() - t
        fkms = solver.last_kernel_ms()
        fbytes = algorithmic_bytes(storage, np.full(len(games), fixed), np.full(len(games), 2))
        out["fixed_300"] = {"value": round(len(games) / fsecs, 3), "unit": "solves/s", "kernel_ms": round(fkms, 3),
                            "roofline_frac": round(fbytes / (fkms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                            "max_expl_pot_pct": round(float((fe / pots).max()) * 100, 4)}
        log(f"fixed 300 iterations: {out['fixed_300']}")
        solver.set_targets(targets)
    # the plain single-subgame path on the same subgames: one game per batch,
    # one launch each (the batched launch is the product; this is its yardstick)
    n_single = min(args.single_sample, len(games))
    if n_single > 0:
        t = time.perf_counter()
        single_kms = 0.0
        for i in range(n_single):
            one = P.GpuBatchSolver(0)
            one.prepare_batch([games[i]], args.iterations, targets[i:i + 1] if args.mode == "solve" else None)
            one.solve(args.iterations, interval)
            single_kms += one.last_kernel_ms()
        single_rate = n_single / (time.perf_counter() - t)
        # per CU: a single subgame runs on one workgroup of one CU, the batch on every CU;
        # device time only on both sides
        cus = solver.launch_info()["num_cus"]
        per_cu = (value / cus) / (n_single / (single_kms * 1e-3))
        out["single_subgame"] = {"value": round(single_rate, 3), "unit": "solves/s", "sample": n_single,
                                 "batched_over_single": round(value / single_rate, 2),
                                 "device_ms_per_game": round(single_kms / n_single, 3),
                                 "batched_per_cu_over_single_device": round(per_cu, 3)}
        log(f"single-subgame path: {single_rate:.1f} solves/s on {n_single} subgames; batched is "
            f"{value / single_rate:.1f}x; per CU, device time only: {per_cu:.2f}x")
    if args.turn_iters > 0:
        out["turn"] = turn_bench(P, W, args.turn_iters, args.seed, cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        n_sample = min(len(subs), 512)
        out["cpu_baseline"] = cpu_baseline(subs[:n_sample], args.iterations, args.mode, interval)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
