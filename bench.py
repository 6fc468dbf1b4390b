#!/usr/bin/env python3
"""Benchmark of the pRRTC hot path on one MI355X.

Workload (one step = one pass of the hot path over one batch): the Panda
edge-validation microbench -- 1M random edges (endpoints uniform in the joint
limits, length <= 0.5 rad) at granularity 128 against 24 oriented cuboids
and 8 vertical cylinders.  Inputs are resident in HBM before timing; each
step is one launch of the edge-validation kernel over all edges.

    value        edge checks / s of the timed steps (edges whose verdict was
                 decided, all of them)
    roofline     the validation kernel on the collision-free subset of the
                 same edges (every pose fully evaluated): reference flops /
                 kernel time, against MI355X's 157.3 TFLOP/s FP32 vector peak
    plan         the configuration's single-problem plan (g=128, dynamic
                 domain) over seeded start/goal pairs: p50/p99 solve ms
    batched      plans/s of a batch of Fetch problems in one persistent launch
                 vs the same problems one launch each (identical results)
    cpu_baseline the CPU oracle on a bounded sample of the same edges,
                 all host threads; paths re-validated on the CPU

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--edges", type=int, default=1_000_000)
    ap.add_argument("--granularity", type=int, default=128)
    ap.add_argument("--plans", type=int, default=16, help="single-problem plans for p50/p99")
    ap.add_argument("--batch", type=int, default=256, help="Fetch problems of the batched vs plain comparison")
    ap.add_argument("--nnc", type=int, default=64, help="num_new_configs of the batched problems")
    ap.add_argument("--cpu-edges", type=int, default=4000)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plans", action="store_true")
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

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps of {args.edges} edges")
    ms = time_launches(step, args.steps, torch)
    free = d_out.cpu().numpy().astype(bool)
    value = args.edges / (ms / 1e3)
    log(f"{ms:.3f} ms/step, {value / 1e6:.2f} M edge checks/s, {free.mean() * 100:.1f}% of edges free")

    # -------------------------------------------------------------- roofline
    free_edges = np.ascontiguousarray(edges[free])
    n_free = len(free_edges)
    reps = max(1, (1 << 17) // max(1, n_free))
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
    # reference flops per pose, from a sample of the free poses
    rng = np.random.default_rng(5)
    pick = roof_edges[rng.integers(0, n_roof, 512)]
    ks = rng.integers(1, g + 1, size=(512, 1)).astype(np.float32)
    poses = (pick[:, 0] + ks * ((pick[:, 1] - pick[:, 0]) / np.float32(g))).astype(np.float32)
    centres = P.fk_spheres("panda", poses)
    per_pose = float(accounting.pose_flops("panda", scene, centres).mean())
    flops = per_pose * n_roof * g
    achieved = flops / (roof_ms / 1e3)
    log(f"free set: {n_roof} edges, {roof_ms:.3f} ms, {per_pose:.0f} flops/pose, {achieved / 1e12:.2f} TFLOP/s")

    out = {
        "metric": "edge checks/sec (Panda, 1M random edges, g=128); plans/sec batched vs one-per-launch; "
                  "p50/p99 solve ms",
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
        "data": "synthetic (seeded scenes and edges)",
        "config": {"workload": "panda edge-validation microbench: 1M random edges <= 0.5 rad, granularity 128, "
                               "24 oriented cuboids + 8 vertical cylinders",
                   "edges": args.edges, "granularity": g, "free_edge_fraction": float(free.mean()),
                   "parallelism": "single GPU"},
        "pose_checks_per_s_free": n_roof * g / (roof_ms / 1e3),
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32_PACKED / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_PACKED,
                     "frac_of_nonpacked_peak": achieved / PEAK_FP32, "traffic": None,
                     "kernel": "k_validate_edges<PandaRobot>", "kernel_ms": roof_ms,
                     "flops_per_pose": per_pose, "poses_per_launch": n_roof * g},
    }

    # ----------------------------------------------------------------- plans
    plan_paths = []
    if not args.skip_plans:
        log(f"single-problem plans: {args.plans}")
        probs = workloads.c1_problems(args.plans, seed=1)
        st = dict(num_new_configs=128, granularity=128, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=1,
                  dd_radius=4.0, dd_alpha=1e-4, dd_min_radius=1.0, max_samples=100000, max_iters=20000)
        P.solve("panda", probs[0]["start"], probs[0]["goals"], scene, dict(st, seed=0))   # warm
        times, solved = [], 0
        for i, p in enumerate(probs):
            t0 = time.perf_counter()
            res, path = P.solve("panda", p["start"], p["goals"], p["scene"], dict(st, seed=i))
            times.append((time.perf_counter() - t0) * 1e3)
            solved += res["solved"]
            if res["solved"]:
                plan_paths.append(("panda", p["scene"], path, 128))
        out["plan"] = {"problems": len(probs), "solved": solved, "p50_ms": float(np.percentile(times, 50)),
                       "p99_ms": float(np.percentile(times, 99)), "settings": "nnc 128, g 128, dd on (r 4, "
                       "alpha 1e-4, min 1), range 0.5, balance 2"}
        log(f"plans: {solved}/{len(probs)} solved, p50 {out['plan']['p50_ms']:.1f} ms, "
            f"p99 {out['plan']['p99_ms']:.1f} ms")

        # batched vs plain
        log(f"batched vs plain: {args.batch} Fetch problems")
        bprobs = workloads.c2_problems(args.batch, seed=0)
        bst = dict(num_new_configs=args.nnc, granularity=64, range=0.5, balance=2, tree_ratio=1.0,
                   dynamic_domain=0, max_samples=100000, max_iters=2000, seed=1)
        P.solve_batch(bprobs[:8], bst)                                                       # warm
        t0 = time.perf_counter()
        bres, bpaths, bstats = P.solve_batch(bprobs, bst)
        t_batch = time.perf_counter() - t0
        t0 = time.perf_counter()
        plain = [P.solve(p["robot"], p["start"], p["goals"], p["scene"], bst) for p in bprobs]
        t_plain = time.perf_counter() - t0
        same = all(r["iters"] == pr[0]["iters"] and r["solved"] == pr[0]["solved"] and
                   pa.tobytes() == pr[1].tobytes() for r, pa, pr in zip(bres, bpaths, plain))
        n_solved = sum(r["solved"] for r in bres)
        out["batched"] = {"problems": len(bprobs), "solved": n_solved,
                          "plans_per_s_batched": n_solved / t_batch,
                          "plans_per_s_plain": sum(r[0]["solved"] for r in plain) / t_plain,
                          "speedup": (n_solved / t_batch) / max(1e-9, sum(r[0]["solved"] for r in plain) / t_plain),
                          "batch_wall_s": t_batch, "plain_wall_s": t_plain,
                          "batch_kernel_ms": bstats["kernel_ns"] / 1e6,
                          "edge_checks_per_s_batched": bstats["edge_checks"] / (bstats["kernel_ns"] / 1e9),
                          "results_identical": bool(same),
                          "settings": f"fetch, g 64, nnc {args.nnc}, range 0.5, balance 2"}
        log(f"batched {out['batched']['plans_per_s_batched']:.1f} plans/s vs plain "
            f"{out['batched']['plans_per_s_plain']:.1f} plans/s, identical={same}")
        for p, r, pa in zip(bprobs, bres, bpaths):
            if r["solved"]:
                plan_paths.append(("fetch", p["scene"], pa, 64))

    # ---------------------------------------------------------- cpu baseline
    if not args.skip_cpu:
        from oracle import oracle
        from prrtc_amd.robots import load_model
        threads = min(16, os.cpu_count() or 1)
        R = oracle.OracleRobot(load_model("panda"))
        sample = np.ascontiguousarray(edges[: args.cpu_edges])
        t0 = time.perf_counter()
        cpu_free = R.validate_edges(scene, sample, g, threads=threads)
        dt = time.perf_counter() - t0
        agree = float((cpu_free == free[: args.cpu_edges]).mean())
        out["cpu_baseline"] = {"value": len(sample) / dt, "unit": "edge checks/s", "cores": threads,
                               "kind": "port",
                               "sample": f"first {len(sample)} of the same edges, g={g}, CPU oracle "
                                         f"(C, -O3), {threads} threads",
                               "verdict_agreement": agree}
        log(f"cpu: {len(sample) / dt:.0f} edge checks/s on {threads} threads, agreement {agree:.4f}")
        # every returned path re-validated edge by edge on the CPU
        models = {}
        bad = 0
        for robot, sc, path, gg in plan_paths:
            if robot not in models:
                models[robot] = oracle.OracleRobot(load_model(robot))
            bad += models[robot].revalidate(sc, path, gg) != -1
        out["paths_revalidated"] = {"paths": len(plan_paths), "valid": len(plan_paths) - bad}
        log(f"paths re-validated on the CPU: {len(plan_paths) - bad}/{len(plan_paths)} valid")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
