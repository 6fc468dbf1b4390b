#!/usr/bin/env python3
"""Benchmark of the pRRTC hot path on one MI355X.

Headline workload (one step = one pass of the hot path over one batch): the
Panda edge-validation microbench -- 1M random edges (endpoints uniform in the
joint limits, length <= 0.5 rad) at granularity 128 against 24 oriented
cuboids and 8 vertical cylinders.  Inputs are resident in HBM before timing;
each step is one launch of the edge-validation kernel over all edges.

    value        edge checks / s of the timed steps (all edges decided)
    roofline     the validation kernel on the collision-free subset of the
                 same edges (every pose fully evaluated): reference flops /
                 kernel time against MI355X's FP32 vector peak
    plan         single-problem plans of the same scene (g=128, dynamic
                 domain): p50/p99 solve ms (the library's clock, API call ->
                 result; the Python call around it beside it)
    c0           plain Panda problem (12 spheres, nnc 128, g 32) over seeds:
                 p50/p99 solve ms (same clock), success, cost; CPU
                 RRT-Connect beside it
    batched      256 Fetch problems (8-64 mixed obstacles, g 64): plans/s of
                 one persistent launch vs one launch per problem, and of the
                 same batch as graph-replayed round grids
    batched_per_robot  the same comparison for Panda and Baxter on the
                 shelf problem sets of the mix (each class's settings), and
                 for UR5, the robot built by the model compiler
    mix          128 Baxter + 256 Panda + 128 Fetch shelf problems (g 128,
                 dynamic domain, balance 1) in one call: one persistent grid
                 serves the three robots
    mbm          every valid problem of the reference's Panda and Fetch MBM
                 problem files with its harness's settings: success rate,
                 cost, solve times; one solve each vs one batch
    serving      Poisson request stream (Panda:Fetch:Baxter 2:1:1, fresh
                 4-128 primitive scenes) through the continuous-batching
                 server: p50/p99 latency, admission wait, plans/s
    cpu_baseline the C oracle on a bounded sample of the same edges on the
                 host's cores; every returned path re-validated on the CPU

Prints one JSON line.  Only N=1 is supported (one GPU).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_PACKED = 157.3e12
PEAK_FP32 = 78.6e12
HBM_PEAK = 8.0e12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_launches(fn, steps, torch):
    """Average device time of `fn` (launches on torch's current stream)."""
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def pct(a, q):
    return float(np.percentile(np.asarray(a, dtype=np.float64), q)) if len(a) else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--edges", type=int, default=1_000_000)
    ap.add_argument("--granularity", type=int, default=128)
    ap.add_argument("--plans", type=int, default=16, help="single-problem plans for p50/p99")
    ap.add_argument("--c0-seeds", type=int, default=128, help="C0-style problems (problem and planner seed i)")
    ap.add_argument("--c0-rounds", type=int, default=20000, help="max_iters of each C0 problem")
    ap.add_argument("--batch-rounds", type=int, default=3000,
                    help="max_iters of every problem of the batched vs plain comparisons (both forms)")
    ap.add_argument("--batch", type=int, default=256, help="Fetch problems of the batched vs plain comparison")
    ap.add_argument("--robot-batch", type=int, default=96,
                    help="Panda and Baxter shelf problems, and UR5 problems, of their batched vs plain comparisons")
    ap.add_argument("--mix-scale", type=float, default=1.0, help="fraction of the 128/256/128 uneven mix")
    ap.add_argument("--serve-rate", type=float, default=500.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-deadline-ms", type=int, default=250, help="per-request deadline of the server (0: none)")
    ap.add_argument("--serve-young", type=int, default=0,
                    help="server young_rounds: rounds a request is served ahead of older ones (0: default 64, "
                         "-1: round-robin)")
    ap.add_argument("--cpu-edges", type=int, default=500_000)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the CPU baseline (at most the CPUs this process may use)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plans", action="store_true")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--skip-mbm", action="store_true", help="skip the full MBM dataset run")
    ap.add_argument("--mbm-batch-samples", type=int, default=100000,
                    help="max_samples of the batched MBM run (tree storage per problem)")
    ap.add_argument("--profile-plan", action="store_true",
                    help="only the C2 Fetch batch (one persistent k_plan launch after one warm-up batch), for "
                         "rocprofv3 runs of the planner kernel")
    ap.add_argument("--profile-run", action="store_true",
                    help="only the free-set validation launches and one 256 MiB device copy (for rocprofv3 "
                         "counter runs: the copy's counters scale FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: independent problems shard across GPUs with no exchange, "
              "and this build measures one GPU's share (--gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch

    import prrtc_amd as P
    from prrtc_amd import accounting, workloads

    P.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream().cuda_stream
    g = args.granularity
    # launch shapes and num_new_configs picked by prrtc_amd.autotune
    from prrtc_amd import autotune
    tuned = autotune.apply()

    if args.profile_plan:
        bprobs = workloads.c2_problems(args.batch, seed=0)
        nnc_b = tuned.get("batch_nnc", {}).get("fetch", 128)
        bst = dict(granularity=64, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0, max_samples=100000,
                   seed=1, num_new_configs=nnc_b, max_iters=args.batch_rounds)
        P.solve_batch(bprobs[:8], dict(bst, max_iters=50))
        res, _, stats = P.solve_batch(bprobs, bst)
        flops = sum(r["pose_checks"] * accounting.gate_pass_flops("fetch", p["scene"]) for p, r in zip(bprobs, res))
        print(json.dumps({"profile_plan": True, "problems": len(bprobs), "solved": stats["solved"],
                          "kernel_ms": stats["kernel_ns"] / 1e6, "rounds": stats["rounds"],
                          "pose_checks": stats["pose_checks"], "edge_checks": stats["edge_checks"],
                          "estimated_flops": flops}), flush=True)
        return

    # ---------------------------------------------------------------- inputs
    log("building the scene and the edges")
    scene = workloads.c1_scene(1)
    packed = P.pack_scene(scene)
    edges = workloads.c1_edges(args.edges, seed=0)
    d_scene = torch.from_numpy(packed).to(dev)
    d_edges = torch.from_numpy(edges).to(dev)
    d_out = torch.zeros(args.edges, dtype=torch.uint8, device=dev)

    def step():
        P.validate_edges_async("panda", d_scene.data_ptr(), packed.size, d_edges.data_ptr(), args.edges, g,
                               d_out.data_ptr(), stream)

    for _ in range(0 if args.profile_run else args.warmup):
        step()
    torch.cuda.synchronize()
    main_steps = 1 if args.profile_run else args.steps   # a profile run needs only the free set
    log(f"timing {main_steps} steps of {args.edges} edges")
    t_wall = time.perf_counter()
    ms = time_launches(step, main_steps, torch)
    t_wall = time.perf_counter() - t_wall
    free = d_out.cpu().numpy().astype(bool)
    value = args.edges / (ms / 1e3)
    log(f"{ms:.3f} ms/step, {value / 1e6:.2f} M edge checks/s, {free.mean() * 100:.1f}% of edges free")

    # -------------------------------------------------------------- roofline
    free_edges = np.ascontiguousarray(edges[free])
    n_free = len(free_edges)
    # the free set repeated (rounded up) and cut to exactly 131,072 edges
    reps = max(1, -(-(1 << 17) // max(1, n_free)))
    roof_edges = np.ascontiguousarray(np.tile(free_edges, (reps, 1, 1))[: 1 << 17]) if n_free else free_edges
    n_roof = len(roof_edges)
    d_roof = torch.from_numpy(roof_edges).to(dev)
    d_roof_out = torch.zeros(n_roof, dtype=torch.uint8, device=dev)

    def roof_step():
        P.validate_edges_async("panda", d_scene.data_ptr(), packed.size, d_roof.data_ptr(), n_roof, g,
                               d_roof_out.data_ptr(), stream)

    roof_step()
    roof_ms = time_launches(roof_step, args.steps, torch)
    assert bool(d_roof_out.all()), "free-edge set must validate free"
    if args.profile_run:
        # the streaming nearest-node case of the bench (one query per tree)
        n_trees, n_nodes, ndof = 256, 100_000, 8
        ncap = ((n_nodes + 1023) // 1024) * 1024
        gen = torch.Generator(device=dev).manual_seed(3)
        d_trees = torch.rand((n_trees, ndof, ncap), generator=gen, device=dev)
        d_q = torch.rand((n_trees, 1, ndof), generator=gen, device=dev)
        d_idx = torch.zeros(n_trees, dtype=torch.int32, device=dev)
        d_d2 = torch.zeros(n_trees, dtype=torch.float32, device=dev)
        nn_ms = time_launches(lambda: P.nearest_async(ndof, d_trees.data_ptr(), ncap, n_nodes, n_trees,
                                                      d_q.data_ptr(), 1, 4, d_idx.data_ptr(), d_d2.data_ptr(),
                                                      stream), args.steps, torch)
        # then the planner-like case: 64 queries per tree (the queries of one
        # tree share its lines through L2)
        qpt = 64
        d_q64 = torch.rand((n_trees, qpt, ndof), generator=gen, device=dev)
        d_idx64 = torch.zeros(n_trees * qpt, dtype=torch.int32, device=dev)
        d_d264 = torch.zeros(n_trees * qpt, dtype=torch.float32, device=dev)
        nn64_ms = time_launches(lambda: P.nearest_async(ndof, d_trees.data_ptr(), ncap, n_nodes, n_trees,
                                                        d_q64.data_ptr(), qpt, 4, d_idx64.data_ptr(),
                                                        d_d264.data_ptr(), stream), args.steps, torch)
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)       # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"profile_run": True, "validate_launches": args.steps + 1, "edges_per_launch": n_roof,
                          "kernel_ms": roof_ms, "nearest_launches": args.steps, "nearest_ms": nn_ms,
                          "nearest_bytes_per_launch": 4 * ndof * n_nodes * n_trees,
                          "nearest64_launches": args.steps, "nearest64_ms": nn64_ms,
                          "copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}), flush=True)
        return
    rng = np.random.default_rng(5)
    pick = roof_edges[rng.integers(0, n_roof, 512)]
    ks = rng.integers(1, g + 1, size=(512, 1)).astype(np.float32)
    poses = (pick[:, 0] + ks * ((pick[:, 1] - pick[:, 0]) / np.float32(g))).astype(np.float32)
    centres = P.fk_spheres("panda", poses)
    per_pose = float(accounting.pose_flops("panda", scene, centres).mean())
    flops = per_pose * n_roof * g
    achieved = flops / (roof_ms / 1e3)
    log(f"free set: {n_roof} edges, {roof_ms:.3f} ms, {per_pose:.0f} flops/pose, {achieved / 1e12:.2f} TFLOP/s")
    traffic = None
    valu = None   # hardware VALU counters of the same kernel from the committed PMC summary
    tpath = os.path.join(ROOT, "profiles", "validate_edges_panda_free_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_launch")
        if tj.get("valu"):
            valu = {k: tj["valu"][k] for k in ("valu_busy", "hw_fp32_tflops", "valu_insts") if k in tj["valu"]}
            valu["source"] = "rocprofv3 --pmc, profiles/validate_edges_panda_c1.md"

    out = {
        "metric": "plans/sec batched vs one-per-launch; p50/p99 solve ms; edge checks/sec",
        "value": value,
        "unit": "edge checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes, edges and problems)",
        "config": {"workload": "panda edge-validation microbench: 1M random edges <= 0.5 rad, granularity 128, "
                               "24 oriented cuboids + 8 vertical cylinders (capsules)",
                   "edges": args.edges, "granularity": g, "free_edge_fraction": float(free.mean()),
                   "parallelism": "single GPU"},
        "timed_wall_s": t_wall,
        "pose_checks_per_s_free": n_roof * g / (roof_ms / 1e3),
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32_PACKED / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_PACKED,
                     "frac_of_nonpacked_peak": achieved / PEAK_FP32, "traffic": traffic,
                     "kernel": "k_validate_edges<PandaRobot, true>", "kernel_ms": roof_ms,
                     "flops_per_pose": per_pose, "poses_per_launch": n_roof * g, "valu_counters": valu},
    }
    paths_to_check = []   # (robot, scene, path, g)

    # ------------------------------------------------- nearest-node search
    # 256 Fetch-sized trees of 100k nodes (SoA, 822 MB: beyond L2 and MALL),
    # 64 queries per tree (one planner round of nnc 64); bytes per query =
    # 4 * dof * n (every coordinate read once)
    n_trees, n_nodes, qpt, ndof = 256, 100_000, 64, 8
    ncap = ((n_nodes + 1023) // 1024) * 1024
    gen = torch.Generator(device=dev).manual_seed(3)
    d_trees = torch.rand((n_trees, ndof, ncap), generator=gen, device=dev)
    d_q = torch.rand((n_trees, qpt, ndof), generator=gen, device=dev)
    d_idx = torch.zeros(n_trees * qpt, dtype=torch.int32, device=dev)
    d_d2 = torch.zeros(n_trees * qpt, dtype=torch.float32, device=dev)
    nn_waves = tuned.get("nearest_waves", 4)

    def nn_step():
        P.nearest_async(ndof, d_trees.data_ptr(), ncap, n_nodes, n_trees, d_q.data_ptr(), qpt, nn_waves,
                        d_idx.data_ptr(), d_d2.data_ptr(), stream)

    nn_step()
    nn_ms = time_launches(nn_step, args.steps, torch)
    nn_bytes = 4.0 * ndof * n_nodes * n_trees * qpt

    # streaming case: one query per tree, so every tree byte comes from HBM once
    def nn_stream():
        P.nearest_async(ndof, d_trees.data_ptr(), ncap, n_nodes, n_trees, d_q.data_ptr(), 1, nn_waves,
                        d_idx.data_ptr(), d_d2.data_ptr(), stream)

    nn_stream()
    st_ms = time_launches(nn_stream, args.steps, torch)
    st_bytes = 4.0 * ndof * n_nodes * n_trees
    out["nearest"] = {"trees": n_trees, "nodes": n_nodes, "dof": ndof, "waves_per_query": nn_waves,
                      "planner_like": {"queries_per_tree": qpt, "kernel_ms": nn_ms,
                                       "algorithmic_GBps": nn_bytes / (nn_ms / 1e3) / 1e9,
                                       "note": "the 64 queries of a tree share its lines through L2"},
                      "streaming": {"queries_per_tree": 1, "kernel_ms": st_ms,
                                    "GBps": st_bytes / (st_ms / 1e3) / 1e9,
                                    "frac_of_hbm_peak": st_bytes / (st_ms / 1e3) / HBM_PEAK},
                      "bytes_per_query": 4 * ndof * n_nodes}
    del d_trees
    log(f"nearest: 64 queries/tree {nn_ms:.3f} ms ({out['nearest']['planner_like']['algorithmic_GBps']:.0f} GB/s "
        f"algorithmic); 1 query/tree {st_ms:.3f} ms = {out['nearest']['streaming']['GBps']:.0f} GB/s "
        f"({out['nearest']['streaming']['frac_of_hbm_peak'] * 100:.1f}% of 8 TB/s)")

    if not args.skip_plans:
        # ------------------------------------------------ plan (same scene)
        log(f"single-problem plans: {args.plans}")
        probs = workloads.c1_problems(args.plans, seed=1)
        st = dict(num_new_configs=128, granularity=128, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=1,
                  dd_radius=4.0, dd_alpha=1e-4, dd_min_radius=1.0, max_samples=100000, max_iters=20000)
        P.solve("panda", probs[0]["start"], probs[0]["goals"], scene, dict(st, seed=0))   # warm
        # solve time = the library's own clock, API call -> result
        # (prrtc_result.wall_ns, the reference's measure, pRRTC.cu:615, 763);
        # the Python call around it (argument packing, ctypes) beside it
        times, py_times, solved = [], [], 0
        for i, p in enumerate(probs):
            t0 = time.perf_counter()
            res, path = P.solve("panda", p["start"], p["goals"], p["scene"], dict(st, seed=i))
            py_times.append((time.perf_counter() - t0) * 1e3)
            times.append(res["wall_ns"] / 1e6)
            solved += res["solved"]
            if res["solved"]:
                paths_to_check.append(("panda", p["scene"], path, 128))
        out["plan"] = {"problems": len(probs), "solved": solved, "p50_ms": pct(times, 50), "p99_ms": pct(times, 99),
                       "python_call_p50_ms": pct(py_times, 50),
                       "settings": "panda, nnc 128, g 128, dd on (r 4, alpha 1e-4, min 1), range 0.5, balance 2"}
        log(f"plans: {solved}/{len(probs)} solved, p50 {out['plan']['p50_ms']:.1f} ms, "
            f"p99 {out['plan']['p99_ms']:.1f} ms")

        # ------------------------------------------------------------- c0
        # C0-style problems, one per seed (a fresh 12-sphere scene, start and
        # goal each): the seed changes the problem, so the figures are
        # distributions over problems, not one plan repeated
        c0_st = dict(num_new_configs=128, granularity=32, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
                     max_samples=100000, max_iters=args.c0_rounds)
        c0_probs = [workloads.c0_problem(s) for s in range(args.c0_seeds)]
        P.solve("panda", c0_probs[0]["start"], c0_probs[0]["goals"], c0_probs[0]["scene"], c0_st)   # warm
        times, py_times, costs, rounds, c0_res = [], [], [], [], []
        for seed, p in enumerate(c0_probs):
            t0 = time.perf_counter()
            res, path = P.solve("panda", p["start"], p["goals"], p["scene"], dict(c0_st, seed=seed))
            py_times.append((time.perf_counter() - t0) * 1e3)
            times.append(res["wall_ns"] / 1e6)   # the library's clock, as for the plans above
            c0_res.append(res)
            if res["solved"]:
                costs.append(res["cost"])
                rounds.append(res["iters"])
                paths_to_check.append(("panda", p["scene"], path, 32))
        solved = len(costs)
        out["c0"] = {"problems": len(c0_probs), "solved": solved, "success_rate": solved / len(c0_probs),
                     "p50_ms": pct(times, 50), "p99_ms": pct(times, 99), "python_call_p50_ms": pct(py_times, 50),
                     "cost_p10": pct(costs, 10), "cost_p50": pct(costs, 50), "cost_p90": pct(costs, 90),
                     "rounds_p50": pct(rounds, 50), "rounds_p90": pct(rounds, 90),
                     "settings": f"panda, 12 spheres, problem and planner seed i for i < {len(c0_probs)}, nnc 128, "
                                 f"g 32, range 0.5, balance 2, ratio 1, dd off, max_iters {args.c0_rounds}"}
        log(f"c0: {solved}/{len(c0_probs)} solved, p50 {out['c0']['p50_ms']:.2f} ms, p99 {out['c0']['p99_ms']:.2f} ms, "
            f"cost p10/p50/p90 {out['c0']['cost_p10']:.2f}/{out['c0']['cost_p50']:.2f}/{out['c0']['cost_p90']:.2f}")

        # -------------------------------------------------- batched vs plain
        # all 256 Fetch problems, both forms under the same round budget
        # (max_iters): a problem that never connects costs its budget in
        # both.  Each form runs the num_new_configs the autotuner picked for it.
        log(f"batched vs plain: {args.batch} Fetch problems, {args.batch_rounds} rounds each at most")
        bprobs = workloads.c2_problems(args.batch, seed=0)
        base = dict(granularity=64, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0, max_samples=100000,
                    seed=1)
        nnc_b = tuned.get("batch_nnc", {}).get("fetch", 128)
        nnc_p = tuned.get("plain_nnc", {}).get("fetch", 256)
        bst = dict(base, num_new_configs=nnc_b, max_iters=args.batch_rounds)
        pst = dict(base, num_new_configs=nnc_p, max_iters=args.batch_rounds)
        P.solve_batch(bprobs[:8], dict(bst, max_iters=50))            # warm
        t0 = time.perf_counter()
        bres, bpaths, bstats = P.solve_batch(bprobs, bst)
        t_batch = time.perf_counter() - t0
        t0 = time.perf_counter()
        plain = [P.solve(p["robot"], p["start"], p["goals"], p["scene"], pst) for p in bprobs]
        t_plain = time.perf_counter() - t0
        # the non-persistent comparator: the same batch as graph-replayed round grids
        t0 = time.perf_counter()
        gres, _, gstats = P.solve_batch(bprobs, bst, graph_rounds=True)
        t_graph = time.perf_counter() - t0
        same_graph = all(r["iters"] == g["iters"] and r["cost"] == g["cost"] for r, g in zip(bres, gres))
        sub = min(32, len(bprobs))
        same = all(r["iters"] == pr[0]["iters"] and pa.tobytes() == pr[1].tobytes() for r, pa, pr in
                   zip(bres[:sub], bpaths[:sub], [P.solve(p["robot"], p["start"], p["goals"], p["scene"], bst)
                                                  for p in bprobs[:sub]]))
        n_b = sum(r["solved"] for r in bres)
        n_p = sum(r[0]["solved"] for r in plain)
        # k_plan: algorithmic work estimate of the batch's validation passes
        # (poses put on lanes x (FK + every gate sphere against every
        # obstacle) per problem) over the persistent kernel's time
        plan_flops = sum(r["pose_checks"] * accounting.gate_pass_flops("fetch", p["scene"])
                         for p, r in zip(bprobs, bres))
        plan_s = bstats["kernel_ns"] / 1e9
        out["batched"] = {"problems": len(bprobs), "round_budget": args.batch_rounds,
                          "solved_batched": n_b, "solved_plain": n_p,
                          "plans_per_s_batched": n_b / t_batch, "plans_per_s_plain": n_p / t_plain,
                          "speedup": (n_b / t_batch) / max(1e-9, n_p / t_plain),
                          "batch_wall_s": t_batch, "plain_wall_s": t_plain,
                          "batch_kernel_ms": bstats["kernel_ns"] / 1e6, "rounds_max": bstats["rounds"],
                          "edge_checks_per_s_batched": bstats["edge_checks"] / plan_s,
                          "pose_checks_per_s_batched": bstats["pose_checks"] / plan_s,
                          "plans_per_s_graph_rounds": sum(r["solved"] for r in gres) / t_graph,
                          "graph_rounds_wall_s": t_graph, "graph_rounds_kernel_ms": gstats["kernel_ns"] / 1e6,
                          "graph_rounds_identical": bool(same_graph),
                          "nnc_batched": nnc_b, "nnc_plain": nnc_p,
                          "cost_p50": pct([r["cost"] for r in bres if r["solved"]], 50),
                          "cost_p90": pct([r["cost"] for r in bres if r["solved"]], 90),
                          "same_settings_identical_results": bool(same),
                          "settings": f"fetch, g 64, range 0.5, balance 2, dd off, max_iters {args.batch_rounds} in "
                                      "both forms; nnc per form from the autotuner"}
        log(f"batched {out['batched']['plans_per_s_batched']:.1f} plans/s vs plain "
            f"{out['batched']['plans_per_s_plain']:.1f} plans/s ({out['batched']['speedup']:.2f}x; {n_b}/{n_p} of "
            f"{len(bprobs)} solved), identical={same}; graph-replayed rounds "
            f"{out['batched']['plans_per_s_graph_rounds']:.1f} plans/s, identical={same_graph}")
        # the problems the batch connected, again without the round budget
        # (max_iters 20000): the ratio of the solvable work alone
        feas = [p for p, r in zip(bprobs, bres) if r["solved"]]
        t0 = time.perf_counter()
        fres, _, fstats = P.solve_batch(feas, dict(bst, max_iters=20000))
        tf_batch = time.perf_counter() - t0
        t0 = time.perf_counter()
        fplain = [P.solve(p["robot"], p["start"], p["goals"], p["scene"], dict(pst, max_iters=20000))
                  for p in feas]
        tf_plain = time.perf_counter() - t0
        nf_b = sum(r["solved"] for r in fres)
        nf_p = sum(r[0]["solved"] for r in fplain)
        out["batched"]["connectable"] = {
            "problems": len(feas), "solved_batched": nf_b, "solved_plain": nf_p,
            "plans_per_s_batched": nf_b / tf_batch, "plans_per_s_plain": nf_p / tf_plain,
            "speedup": (nf_b / tf_batch) / max(1e-9, nf_p / tf_plain), "batch_kernel_ms": fstats["kernel_ns"] / 1e6,
            "hardest_alone_ms": max(r[0]["kernel_ns"] for r in fplain) / 1e6,
            "settings": "the problems the budgeted batch solved, max_iters 20000 in both forms"}
        log(f"connectable subset: {nf_b / tf_batch:.1f} vs {nf_p / tf_plain:.1f} plans/s "
            f"({out['batched']['connectable']['speedup']:.2f}x) on {len(feas)} problems; hardest alone "
            f"{out['batched']['connectable']['hardest_alone_ms']:.1f} ms, batch {fstats['kernel_ns'] / 1e6:.1f} ms")
        out["k_plan_roofline"] = {
            "kernel": "k_plan<FetchRobot>", "workload": f"the {len(bprobs)}-problem Fetch batch above",
            "bound": "valu", "achieved": plan_flops / plan_s / 1e12, "peak": PEAK_FP32_PACKED / 1e12,
            "unit": "TFLOP/s", "frac": plan_flops / plan_s / PEAK_FP32_PACKED,
            "frac_of_nonpacked_peak": plan_flops / plan_s / PEAK_FP32, "kernel_ms": plan_s * 1e3,
            "pose_checks": bstats["pose_checks"], "flops_per_pose_mean": plan_flops / max(1, bstats["pose_checks"]),
            "flops_definition": "per pose put on a lane: FK + every env gate sphere against every obstacle of the "
                                "problem's scene (leaves, self pairs and early exits not counted)",
            "valu_counters": None}
        kpath = os.path.join(ROOT, "profiles", "planner_fetch_c2_counters.json")
        if os.path.exists(kpath):
            with open(kpath) as f:
                out["k_plan_roofline"]["valu_counters"] = json.load(f).get("summary")
        log(f"k_plan: {bstats['pose_checks']:.3e} poses, {out['k_plan_roofline']['achieved']:.2f} TFLOP/s "
            f"estimated ({out['k_plan_roofline']['frac'] * 100:.1f}% of packed peak)")

        # ------------------------ batched vs plain, Panda, Baxter and UR5
        # The same comparison on the shelf problem sets of the uneven mix,
        # and on UR5 problems (the robot the model compiler built), one class
        # at a time, with each class's settings and tuned nnc:
        # on the problems a pilot batch connects (max_iters 20000 in both
        # forms), and on every problem under the round budget in both forms.
        from prrtc_amd.autotune import SETTINGS as CLASS_SETTINGS, problems as class_problems
        out["batched_per_robot"] = {}
        for robot in ("panda", "baxter", "ur5"):
            rp = class_problems(robot, args.robot_batch)
            cbase = CLASS_SETTINGS[robot]
            nb_nnc = tuned.get("batch_nnc", {}).get(robot, 128)
            np_nnc = tuned.get("plain_nnc", {}).get(robot, 128)
            row = {"problems": len(rp), "nnc_batched": nb_nnc, "nnc_plain": np_nnc,
                   "granularity": cbase["granularity"]}
            pilot, _, _ = P.solve_batch(rp, dict(cbase, num_new_configs=64, max_iters=2000))
            rf = [p for p, r in zip(rp, pilot) if r["solved"]]
            for key, probs_, rounds in (("connectable", rf, 20000), ("all_budgeted", rp, args.batch_rounds)):
                t0 = time.perf_counter()
                rres, rpaths, _ = P.solve_batch(probs_, dict(cbase, num_new_configs=nb_nnc, max_iters=rounds))
                tb = time.perf_counter() - t0
                t0 = time.perf_counter()
                rplain = [P.solve(p["robot"], p["start"], p["goals"], p["scene"],
                                  dict(cbase, num_new_configs=np_nnc, max_iters=rounds)) for p in probs_]
                tp = time.perf_counter() - t0
                nb_r = sum(r["solved"] for r in rres)
                np_r = sum(r[0]["solved"] for r in rplain)
                row[key] = {"problems": len(probs_), "max_iters": rounds, "solved_batched": nb_r,
                            "solved_plain": np_r, "plans_per_s_batched": nb_r / tb,
                            "plans_per_s_plain": np_r / tp, "speedup": (nb_r / tb) / max(1e-9, np_r / tp)}
                log(f"{robot} {key}: batched {nb_r / tb:.1f} plans/s vs plain {np_r / tp:.1f} plans/s "
                    f"({row[key]['speedup']:.2f}x) on {len(probs_)} problems")
                if key == "connectable":
                    for p, r, pa in zip(probs_, rres, rpaths):
                        if r["solved"]:
                            paths_to_check.append((robot, p["scene"], pa, cbase["granularity"]))
            row["speedup"] = row["connectable"]["speedup"]
            out["batched_per_robot"][robot] = row

        for p, r, pa in zip(bprobs, bres, bpaths):
            if r["solved"]:
                paths_to_check.append(("fetch", p["scene"], pa, 64))

        # -------------------------------------------------------------- mix
        nb, npd, nf = (max(1, int(round(k * args.mix_scale))) for k in (128, 256, 128))
        log(f"uneven mix: {nb} Baxter + {npd} Panda + {nf} Fetch")
        mprobs = workloads.c3_problems(nb, npd, nf, seed=0)
        # nnc 128 for every class: a problem that ends at max_iters costs
        # max_iters x nnc tasks, and the mix has such problems.  The plain
        # form runs the same settings, one problem per launch, so both forms
        # must return the same results bit for bit.
        mst = [dict(num_new_configs=128, granularity=128, range=0.5, balance=1, tree_ratio=0.5,
                    dynamic_domain=1, dd_radius=4.0, dd_alpha=1e-4, dd_min_radius=1.0, max_samples=100000,
                    max_iters=3000, seed=1) for p in mprobs]
        t0 = time.perf_counter()
        mres, mpaths, mstats = P.solve_batch(mprobs, mst)
        t_mix = time.perf_counter() - t0
        t0 = time.perf_counter()
        mplain = [P.solve(p["robot"], p["start"], p["goals"], p["scene"], st) for p, st in zip(mprobs, mst)]
        t_mix_plain = time.perf_counter() - t0
        m_same = all(r["iters"] == r1["iters"] and r["start_tree_size"] == r1["start_tree_size"] and
                     pa.tobytes() == pa1.tobytes() for r, pa, (r1, pa1) in zip(mres, mpaths, mplain))
        n_m = sum(r["solved"] for r in mres)
        n_mp = sum(r[0]["solved"] for r in mplain)
        per_robot = {}
        for p, r in zip(mprobs, mres):
            d = per_robot.setdefault(p["robot"], {"problems": 0, "solved": 0})
            d["problems"] += 1
            d["solved"] += r["solved"]
        out["mix"] = {"problems": len(mprobs), "solved": n_m, "plans_per_s": n_m / t_mix, "wall_s": t_mix,
                      "kernel_ms": mstats["kernel_ns"] / 1e6, "per_robot": per_robot,
                      "solved_plain": n_mp, "plans_per_s_plain": n_mp / t_mix_plain, "plain_wall_s": t_mix_plain,
                      "speedup": (n_m / t_mix) / max(1e-9, n_mp / t_mix_plain), "identical_results": bool(m_same),
                      "settings": "nnc 128, g 128, balance 1, ratio 0.5, dd on, range 0.5, max_iters 3000; one persistent "
                                  "grid for the three robots; plain: the same settings, one problem per launch"}
        log(f"mix: {n_m}/{len(mprobs)} solved in {t_mix:.3f} s ({n_m / t_mix:.1f} plans/s) vs plain "
            f"{n_mp / t_mix_plain:.1f} plans/s ({out['mix']['speedup']:.2f}x), identical={m_same}")
        for p, r, pa in zip(mprobs, mres, mpaths):
            if r["solved"]:
                paths_to_check.append((p["robot"], p["scene"], pa, 128))

    if not args.skip_mbm:
        # ----------------------------------------------------- MBM dataset
        # Every valid problem of the reference's Panda and Fetch problem files
        # (tests/golden/mbm_*_valid.json.gz) with the reference harness's
        # settings (scripts/evaluate_mbm.cpp:154-166, max_iters and
        # max_samples left at 1,000,000): one solve each, as the harness runs
        # them, then all of them in one batch.  The batch lowers max_samples
        # to --mbm-batch-samples (tree storage of ~700 problems); that changes
        # no result unless a tree reaches the limit, which the row-by-row
        # comparison would show.
        from prrtc_amd import mbm
        out["mbm"] = {}
        for robot in ("panda", "fetch"):
            data = mbm.load(os.path.join(ROOT, "tests", "golden", f"mbm_{robot}_valid.json.gz"))
            mst_ = dict(mbm.EVALUATE_SETTINGS)
            log(f"mbm {robot}: every valid problem, one solve each, then one batch")
            t0 = time.perf_counter()
            rows = mbm.evaluate(data, mst_)
            t_plain = time.perf_counter() - t0
            t0 = time.perf_counter()
            brows = mbm.evaluate(data, dict(mst_, max_samples=args.mbm_batch_samples), batch=True)
            t_batch = time.perf_counter() - t0
            keys = ("solved", "iters", "cost", "start_tree_size", "goal_tree_size")
            same = all(all(a[2][k] == b[2][k] for k in keys) and a[3].tobytes() == b[3].tobytes()
                       for a, b in zip(rows, brows))
            sm = mbm.summary(rows)
            n_p = sm["all"]["solved"]
            n_b = sum(r["solved"] for _, _, r, _ in brows)
            out["mbm"][robot] = {
                **sm["all"], "success_rate_per_family": {f: v["success_rate"] for f, v in sm["families"].items()},
                "max_iters_seen": max(r["iters"] for _, _, r, _ in rows),
                "plans_per_s_plain": n_p / t_plain, "plans_per_s_batched": n_b / t_batch,
                "speedup": (n_b / t_batch) / max(1e-9, n_p / t_plain), "plain_wall_s": t_plain,
                "batch_wall_s": t_batch, "solved_batched": n_b, "identical_results": bool(same),
                "settings": "the reference harness's (nnc 512, g 32, range 0.5, balance 2, ratio 1, dd on r 4 "
                            f"alpha 1e-4 min 1, max_iters 1e6); batch max_samples {args.mbm_batch_samples}"}
            log(f"mbm {robot}: {n_p}/{len(rows)} solved, wall p50 {sm['all']['wall_ms']['p50']:.2f} ms p99 "
                f"{sm['all']['wall_ms']['p99']:.2f} ms, cost p50 {sm['all']['cost']['p50']:.2f}; batch "
                f"{out['mbm'][robot]['speedup']:.2f}x, identical={same}")
            for (fam, idx, r, pa), (_, _, p) in zip(rows, mbm.valid_problems(data)):
                if r["solved"]:
                    paths_to_check.append((robot, p["scene"], pa, mst_["granularity"]))

    if not args.skip_serving:
        # ---------------------------------------------------------- serving
        from prrtc_amd import loadgen
        n_req = max(1, int(args.serve_rate * args.serve_seconds))
        log(f"serving: {n_req} requests at {args.serve_rate:.0f}/s")
        reqs = loadgen.request_pool(n_req)
        sv = loadgen.run_stream(reqs, args.serve_rate, args.serve_seconds, young_rounds=args.serve_young,
                                deadline_ms=args.serve_deadline_ms)
        outs = sv.pop("_outs")
        out["serving"] = sv
        log(f"serving: {sv['solved']}/{n_req} solved, p50 {sv['p50_ms']:.1f} ms, "
            f"p99 {sv['p99_ms']:.1f} ms, {sv['plans_per_s']:.1f} plans/s")
        for rq, r, pa, _ in outs[:200]:
            if r["solved"]:
                paths_to_check.append((rq["robot"], rq["scene"], pa, 64))

    # ---------------------------------------------------------- cpu baseline
    if not args.skip_cpu:
        from oracle import oracle
        from prrtc_amd.robots import load_model
        # the host's cores: nproc, and the ones this process may run on; a
        # GPU box gives one GPU's job a 16-thread share of its CPUs
        nproc = os.cpu_count() or 1
        affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
        threads = max(1, min(args.cpu_threads, affinity))
        build = oracle.use_fastest_build()
        host_cpu = ""
        try:
            with open("/proc/cpuinfo") as f:
                host_cpu = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
        except OSError:
            pass
        models = {}

        def model(robot):
            if robot not in models:
                models[robot] = oracle.OracleRobot(load_model(robot))
            return models[robot]

        sample = np.ascontiguousarray(edges[: args.cpu_edges])
        t0 = time.perf_counter()
        cpu_free = model("panda").validate_edges(scene, sample, g, threads=threads)
        dt = time.perf_counter() - t0
        gpu_free = free[: args.cpu_edges]
        agree = float((cpu_free == gpu_free).mean())
        # verdict flips against the reference-rounding checker: each must have
        # a pose within 1e-6 m of contact (distance from the oracle's exhaustive
        # clearance over the edge's g poses)
        flips = np.nonzero(cpu_free != gpu_free)[0]
        near = 0
        if len(flips):
            fe = sample[flips]
            delta = (fe[:, 1] - fe[:, 0]) / np.float32(g)
            fq = (fe[:, 0][:, None, :] + np.arange(1, g + 1, dtype=np.float32)[None, :, None] * delta[:, None, :])
            _, fm = model("panda").check(scene, fq.reshape(-1, fe.shape[-1]).astype(np.float32), margin=True)
            near = int((fm.reshape(len(flips), g).min(axis=1) < 1e-6).sum())
        out["cpu_baseline"] = {"value": len(sample) / dt, "unit": "edge checks/s", "cores": threads, "kind": "port",
                               "sample": f"first {len(sample)} of the same 1M edges, g={g}, C oracle (-O3, "
                                         f"-march={build}), {threads} threads, static partition",
                               "seconds": dt, "verdict_agreement": agree, "host_nproc": nproc,
                               "host_cpus_available": affinity, "host_cpu": host_cpu}
        out["parity"] = {"edges_compared": len(sample), "verdict_flips": int(len(flips)),
                         "near_contact_flips": near, "flips_away_from_contact": int(len(flips)) - near,
                         "checker": "oracle in the reference's rounding (separate multiplies and adds)"}
        log(f"cpu: {len(sample) / dt:.0f} edge checks/s on {threads} threads ({build}; nproc {nproc}, "
            f"{affinity} available), agreement {agree:.6f}, {len(flips)} flips ({near} near contact)")
        one = np.ascontiguousarray(edges[: max(1, args.cpu_edges // 25)])
        t0 = time.perf_counter()
        model("panda").validate_edges(scene, one, g, threads=1)
        dt1 = time.perf_counter() - t0
        out["cpu_baseline_1thread"] = {"value": len(one) / dt1, "unit": "edge checks/s", "cores": 1,
                                       "sample": f"first {len(one)} of the same edges", "seconds": dt1}
        log(f"cpu: {len(one) / dt1:.0f} edge checks/s on 1 thread")
        from concurrent.futures import ThreadPoolExecutor
        if "c0" in out:
            # the C0 problems through the CPU RRT-Connect, and through the
            # oracle's planner in its sequential mode (each slot sees every
            # earlier insertion at once, the extreme of the reference's
            # concurrent visibility) against the GPU's round-synchronous plans
            pm = model("panda")

            def cpu_c0(ip):
                i, p = ip
                t0 = time.perf_counter()
                res = pm.rrt_connect(p["scene"], p["start"], p["goals"][0], 32, 0.5, 100000, i)[1]
                return (time.perf_counter() - t0) * 1e3, bool(res.solved), float(res.cost)

            def seq_c0(ip):
                i, p = ip
                st = dict(c0_st, seed=i, max_iters=min(args.c0_rounds, 2000))
                res = pm.plan(p["scene"], p["start"], p["goals"], st, mode=1)[1]
                return bool(res.solved), float(res.cost), int(res.iters)

            with ThreadPoolExecutor(threads) as ex:
                rc_out = list(ex.map(cpu_c0, enumerate(c0_probs)))
                seq_out = list(ex.map(seq_c0, enumerate(c0_probs)))
            out["cpu_rrt_connect_c0"] = {"problems": len(c0_probs), "solved": sum(s for _, s, _ in rc_out),
                                         "p50_ms": pct([t for t, _, _ in rc_out], 50),
                                         "p99_ms": pct([t for t, _, _ in rc_out], 99),
                                         "cost_p50": pct([c for _, s, c in rc_out if s], 50), "cores": 1,
                                         "kind": "project CPU RRT-Connect (vamp unavailable offline), one problem "
                                                 "per thread, each timed alone"}
            gc = [r["cost"] for r in c0_res if r["solved"]]
            sc_ = [c for s, c, _ in seq_out if s]
            gi = [r["iters"] for r in c0_res if r["solved"]]
            si = [it for s, _, it in seq_out if s]
            dist = {"problems": len(c0_probs), "gpu_solved": len(gc), "sequential_solved": len(sc_),
                    "gpu_cost_p10_p50_p90": [pct(gc, 10), pct(gc, 50), pct(gc, 90)],
                    "sequential_cost_p10_p50_p90": [pct(sc_, 10), pct(sc_, 50), pct(sc_, 90)],
                    "gpu_rounds_p50_p90": [pct(gi, 50), pct(gi, 90)],
                    "sequential_rounds_p50_p90": [pct(si, 50), pct(si, 90)],
                    "paired_cost_ratio_p50": pct([r["cost"] / c for r, (s, c, _) in zip(c0_res, seq_out)
                                                  if r["solved"] and s and c > 0], 50),
                    "reference_semantics": "oracle or_plan mode 1 (sequential visibility), same settings "
                                           "(max_iters capped at 2000 on the CPU)"}
            try:
                from scipy import stats as sst
                if len(gc) > 1 and len(sc_) > 1:
                    dist["cost_mannwhitney_p"] = float(sst.mannwhitneyu(gc, sc_, alternative="two-sided").pvalue)
                    dist["cost_ks_p"] = float(sst.ks_2samp(gc, sc_).pvalue)
                    dist["rounds_mannwhitney_p"] = float(sst.mannwhitneyu(gi, si, alternative="two-sided").pvalue)
                dist["success_fisher_p"] = float(sst.fisher_exact(
                    [[len(gc), len(c0_probs) - len(gc)], [len(sc_), len(c0_probs) - len(sc_)]])[1])
                dist["test"] = ("two-sided Mann-Whitney U and Kolmogorov-Smirnov on the solved problems' costs "
                                "and rounds; Fisher's exact test on success counts")
            except ImportError:
                dist["test"] = "scipy unavailable"
            out["c0_vs_sequential"] = dist
            log(f"c0 distributions: GPU cost p50 {dist['gpu_cost_p10_p50_p90'][1]:.3f} vs sequential "
                f"{dist['sequential_cost_p10_p50_p90'][1]:.3f} (MW p {dist.get('cost_mannwhitney_p', float('nan')):.3f}), "
                f"solved {len(gc)} vs {len(sc_)}")
        if "batched" in out:
            # the Fetch problems the batch connected through the CPU
            # RRT-Connect, one problem per host thread (ctypes releases the
            # GIL), bounded at 5000 samples per problem
            fprobs = feas
            fr = model("fetch")

            def cpu_solve(p):
                res = fr.rrt_connect(p["scene"], p["start"], p["goals"][0], 64, 0.5, 5000, 1)[1]
                return bool(res.solved), float(res.cost)

            t0 = time.perf_counter()
            with ThreadPoolExecutor(threads) as ex:
                cpu_res = list(ex.map(cpu_solve, fprobs))
            dt = time.perf_counter() - t0
            cpu_solved = sum(sv for sv, _ in cpu_res)
            # path cost against the GPU planner's on the problems both solved
            fgpu = [r for r in bres if r["solved"]]
            both = [(gr["cost"], c) for gr, (sv, c) in zip(fgpu, cpu_res) if sv]
            out["cpu_rrt_connect_c2"] = {"problems": len(fprobs), "solved": int(cpu_solved),
                                         "cost_p50": pct([c for sv, c in cpu_res if sv], 50),
                                         "cost_p90": pct([c for sv, c in cpu_res if sv], 90),
                                         "gpu_over_cpu_cost_p50": pct([a / c for a, c in both if c > 0], 50),
                                         "both_solved": len(both),
                                         "plans_per_s": cpu_solved / dt, "wall_s": dt, "cores": threads,
                                         "kind": "project CPU RRT-Connect (vamp unavailable offline)",
                                         "settings": "fetch, g 64, range 0.5, max 5000 samples per problem"}
            log(f"cpu rrt-connect c2: {cpu_solved}/{len(fprobs)} solved in {dt:.1f} s on {threads} threads; "
                f"path cost p50 {out['cpu_rrt_connect_c2']['cost_p50']:.2f} vs GPU batch "
                f"{out['batched']['cost_p50']:.2f} (GPU/CPU p50 {out['cpu_rrt_connect_c2']['gpu_over_cpu_cost_p50']:.2f} "
                f"on {len(both)} problems both solved)")
        if "mix" in out:
            # the uneven mix through the CPU RRT-Connect (its own robot per
            # problem), bounded at 5000 samples per problem
            from concurrent.futures import ThreadPoolExecutor

            def cpu_mix(p):
                return model(p["robot"]).rrt_connect(p["scene"], p["start"], p["goals"][0], 128, 0.5, 5000,
                                                     1)[1].solved

            for r in sorted({p["robot"] for p in mprobs}):
                model(r)   # built before the worker threads share the table
            t0 = time.perf_counter()
            with ThreadPoolExecutor(threads) as ex:
                mix_solved = list(ex.map(cpu_mix, mprobs))
            dt = time.perf_counter() - t0
            per = {}
            for p, sv in zip(mprobs, mix_solved):
                d = per.setdefault(p["robot"], {"problems": 0, "solved": 0})
                d["problems"] += 1
                d["solved"] += int(sv)
            out["cpu_rrt_connect_c3"] = {"problems": len(mprobs), "solved": int(sum(mix_solved)),
                                         "plans_per_s": sum(mix_solved) / dt, "wall_s": dt, "cores": threads,
                                         "per_robot": per,
                                         "kind": "project CPU RRT-Connect (vamp unavailable offline)",
                                         "settings": "g 128, range 0.5, max 5000 samples per problem"}
            log(f"cpu rrt-connect c3: {sum(mix_solved)}/{len(mprobs)} solved in {dt:.1f} s on {threads} threads")
        bad = 0
        for robot, sc, path, gg in paths_to_check:
            bad += model(robot).revalidate(sc, path, gg) != -1
        out["paths_revalidated"] = {"paths": len(paths_to_check), "valid": len(paths_to_check) - bad}
        log(f"paths re-validated on the CPU: {len(paths_to_check) - bad}/{len(paths_to_check)} valid")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
