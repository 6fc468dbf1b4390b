"""Autotuner of the planner's launch shapes and num_new_configs per robot class.

    python -m prrtc_amd.autotune [--robots fetch,panda,baxter] [--problems 96] [--out PATH]

For every robot class it builds a seeded problem set of the benchmark's kind
(Fetch: 8-64 mixed obstacles, g 64; Panda and Baxter: shelves with 4-128
obstacles, g 128, dynamic domain; UR5: 12 mixed obstacles in reach, g 32,
dynamic domain), keeps the problems a pilot run connects,
and measures on the GPU:

* batched form (one persistent launch): plans/s over num_new_configs x waves
  per task x waves per CU; the best sets ``batch_nnc``, ``waves_per_cu`` and
  ``batch_waves_per_task``;
* plain form (one problem per launch): plans/s over num_new_configs x waves
  per task on a subset; the best sets ``plain_nnc`` and ``waves_per_task``;
* crowd threshold (the running-problem count above which a workgroup's
  waves take one slot each): plans/s of the whole problem set under a round
  budget (--crowd-rounds), where problems that never connect keep long
  rounds running; a count replaces the default only when it is at least
  3% faster.  ``--crowd-only`` runs this stage alone on an existing file.

A lane always carries one pose (64 poses per wavefront); the tuned shape is
how many wavefronts share one task (they split the obstacle tables of each
validation pass) and how many wavefronts a CU holds.  Launch shapes never
change results; num_new_configs does (it is a planner setting), so every
candidate is timed on the same problems and the choice is by solved plans
per second.  The result is JSON, read by bench.py:

    {"launch": {robot: {"waves_per_cu", "waves_per_task", "batch_waves_per_task", "crowd_above"}},
     "batch_nnc": {robot: n}, "plain_nnc": {robot: n},
     "measured": {...every candidate's plans/s...}}

The nnc choice is measured on problems a pilot run found feasible.  A
workload with problems that end at max_iters pays max_iters x nnc tasks for
each of them, so a larger nnc costs such workloads (a serving stream, a
mixed batch with hard scenes) more than it gains; they keep their own nnc.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tuned.json")

SETTINGS = {
    "fetch": dict(granularity=64, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0, max_samples=100000,
                  seed=1),
    "panda": dict(granularity=128, range=0.5, balance=1, tree_ratio=0.5, dynamic_domain=1, dd_radius=4.0,
                  dd_alpha=1e-4, dd_min_radius=1.0, max_samples=100000, seed=1),
    "ur5": dict(granularity=32, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=1, dd_radius=4.0,
                dd_alpha=1e-4, dd_min_radius=1.0, max_samples=100000, seed=1),
    "baxter": dict(granularity=128, range=0.5, balance=1, tree_ratio=0.5, dynamic_domain=1, dd_radius=4.0,
                   dd_alpha=1e-4, dd_min_radius=1.0, max_samples=100000, seed=1),
}


def apply(path: str = DEFAULT_OUT) -> dict:
    """Set every robot's launch shapes from a tuned file (prrtc_set_tuning);
    returns the file's contents, {} when there is no file."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        tuned = json.load(f)
    import prrtc_amd as P
    for robot, t in tuned.get("launch", {}).items():
        P.set_tuning(robot, t.get("waves_per_cu", 0), t.get("waves_per_task", 0), t.get("batch_waves_per_task", 0),
                     t.get("crowd_above", 0))
    return tuned


def log(msg):
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


def problems(robot, n):
    from . import workloads
    if robot == "fetch":
        return workloads.c2_problems(n, seed=0)
    if robot == "ur5":
        return workloads.ur5_problems(n, seed=0)
    counts = {"panda": (0, n, 0), "baxter": (n, 0, 0)}[robot]
    return workloads.c3_problems(*counts, seed=0)


def timed(fn):
    t0 = time.perf_counter()
    out = fn()
    return time.perf_counter() - t0, out


def tune_robot(P, robot, n, pilot_iters, max_iters, plain_n):
    base = SETTINGS[robot]
    probs = problems(robot, n)
    _, (pres, _, _) = timed(lambda: P.solve_batch(probs, dict(base, num_new_configs=64, max_iters=pilot_iters)))
    feas = [p for p, r in zip(probs, pres) if r["solved"]]
    log(f"{robot}: {len(feas)}/{len(probs)} problems connect within {pilot_iters} rounds")
    measured = {"batch": [], "plain": []}
    best_b = None
    for waves in (1, 2, 4):
        for wpc in (4, 8):
            if wpc < waves:
                continue
            # crowd claims (one slot per wave while many problems run) need
            # several waves per workgroup; -1 turns them off, 0 is the default
            for crowd in ((-1, 0) if waves > 1 else (-1,)):
                P.set_tuning(robot, waves_per_cu=wpc, waves_per_task=waves, batch_waves_per_task=waves,
                             crowd_above=crowd)
                for nnc in (32, 64, 128, 256):
                    st = dict(base, num_new_configs=nnc, max_iters=max_iters)
                    dt, (res, _, _) = timed(lambda: P.solve_batch(feas, st))
                    rate = sum(r["solved"] for r in res) / dt
                    measured["batch"].append({"waves_per_task": waves, "waves_per_cu": wpc, "crowd_above": crowd,
                                              "nnc": nnc, "plans_per_s": rate, "wall_s": dt})
                    log(f"{robot} batch waves {waves} waves/CU {wpc} crowd {crowd} nnc {nnc}: {rate:.1f} plans/s")
                    if best_b is None or rate > best_b[0]:
                        best_b = (rate, waves, wpc, nnc, crowd)
    sub = feas[:plain_n]
    best_p = None
    for waves in (2, 4):
        P.set_tuning(robot, waves_per_cu=best_b[2], waves_per_task=waves, batch_waves_per_task=waves)
        for nnc in (64, 128, 256, 512):
            st = dict(base, num_new_configs=nnc, max_iters=max_iters)
            dt, res = timed(lambda: [P.solve(p["robot"], p["start"], p["goals"], p["scene"], st) for p in sub])
            rate = sum(r[0]["solved"] for r in res) / dt
            measured["plain"].append({"waves_per_task": waves, "nnc": nnc, "plans_per_s": rate, "wall_s": dt})
            log(f"{robot} plain waves {waves} nnc {nnc}: {rate:.1f} plans/s")
            if best_p is None or rate > best_p[0]:
                best_p = (rate, waves, nnc)
    P.set_tuning(robot, 0, 0, 0)
    launch = {"waves_per_cu": best_b[2], "waves_per_task": best_p[1], "batch_waves_per_task": best_b[1],
              "crowd_above": best_b[4]}
    return launch, best_b[3], best_p[2], measured


CROWD_CANDIDATES = (0, 8, 16, 32)


def tune_crowd(P, robot, n, launch, nnc, rounds):
    """Crowd threshold of a tuned launch shape on every problem of the set
    (no feasibility filter) under `rounds` rounds: 0 (the default) unless a
    count is at least 3% faster.  Returns (crowd_above, measurements)."""
    if launch["batch_waves_per_task"] < 2 or launch["crowd_above"] < 0:
        return launch["crowd_above"], []
    probs = problems(robot, n)
    st = dict(SETTINGS[robot], num_new_configs=nnc, max_iters=rounds)
    P.set_tuning(robot, launch["waves_per_cu"], launch["waves_per_task"], launch["batch_waves_per_task"], 0)
    P.solve_batch(probs[:8], dict(st, max_iters=50))
    rows, rate = [], {}
    for crowd in CROWD_CANDIDATES:
        P.set_tuning(robot, launch["waves_per_cu"], launch["waves_per_task"], launch["batch_waves_per_task"], crowd)
        dt, (res, _, _) = timed(lambda: P.solve_batch(probs, st))
        rate[crowd] = sum(r["solved"] for r in res) / dt
        rows.append({"crowd_above": crowd, "problems": len(probs), "max_iters": rounds, "plans_per_s": rate[crowd],
                     "wall_s": dt})
        log(f"{robot} crowd {crowd}: {rate[crowd]:.1f} plans/s on {len(probs)} problems, {rounds} rounds at most")
    P.set_tuning(robot, 0, 0, 0)
    best = max(rate, key=rate.get)
    return (best if rate[best] >= 1.03 * rate[0] else 0), rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--robots", default="fetch,panda,baxter")
    ap.add_argument("--problems", type=int, default=96)
    ap.add_argument("--plain-problems", type=int, default=24)
    ap.add_argument("--pilot-iters", type=int, default=2000)
    ap.add_argument("--max-iters", type=int, default=20000)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--crowd-rounds", type=int, default=3000)
    ap.add_argument("--crowd-only", action="store_true",
                    help="run only the crowd-threshold stage, on the launch shapes of --out")
    args = ap.parse_args(argv)
    import prrtc_amd as P
    P.load()
    if args.crowd_only:
        with open(args.out) as f:
            out = json.load(f)
        for robot in args.robots.split(","):
            crowd, rows = tune_crowd(P, robot, args.problems, out["launch"][robot], out["batch_nnc"][robot],
                                     args.crowd_rounds)
            out["launch"][robot]["crowd_above"] = crowd
            out["measured"].setdefault(robot, {})["crowd"] = rows
            log(f"{robot}: crowd_above {crowd}")
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
        log(f"wrote {args.out}")
        return
    out = {"launch": {}, "batch_nnc": {}, "plain_nnc": {}, "measured": {},
           "problems": args.problems}
    for robot in args.robots.split(","):
        n = args.problems if robot != "baxter" else max(8, args.problems // 3)
        launch, nb, npl, measured = tune_robot(P, robot, n, args.pilot_iters, args.max_iters, args.plain_problems)
        # the threshold is a count of running problems: measured on the full
        # set size for every robot
        launch["crowd_above"], measured["crowd"] = tune_crowd(P, robot, args.problems, launch, nb, args.crowd_rounds)
        out["launch"][robot] = launch
        out["batch_nnc"][robot] = nb
        out["plain_nnc"][robot] = npl
        out["measured"][robot] = measured
        log(f"{robot}: launch {launch}, batch nnc {nb}, plain nnc {npl}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()
