"""Motion-benchmark (MBM) problem sets through the planner: the counterpart of
the reference's harness (scripts/evaluate_mbm.cpp, scripts/single_mbm.cpp).

    python -m prrtc_amd.mbm evaluate <robot> <run_name> --problems FILE [--batch] [--out-dir DIR]
                                     [--max-iters N] [--max-samples N] [--summary]
    python -m prrtc_amd.mbm single <robot> <problem_name> <problem_idx> --problems FILE

A problem file is the reference's JSON ({"robot", "joints", "problems":
{family: [problem, ...]}}, each problem with "sphere", "cylinder", "box",
"start", "goals", "valid").  Scenes are built as the reference's loader
builds them (evaluate_mbm.cpp:13-86): spheres as spheres; cylinders as
capsules from centre, Euler angles, radius and length; boxes as oriented
cuboids from centre, Euler angles and half extents -- except in the "box"
family, whose cylinders become cuboids with half extents (r, r, r/2)
(evaluate_mbm.cpp:27-39).  Problems marked invalid are skipped.

``evaluate`` writes <out-dir>/<robot>_<run_name>.csv with the reference's
columns (evaluate_mbm.cpp:88-115); with --batch every valid problem of the
file is solved in one prrtc_solve_batch call instead of one solve each.
--max-iters and --max-samples replace the harness's defaults (1,000,000
each; pRRTC_settings.hh:4-5); --summary prints success rate, path cost and
solve times per family as JSON.  A problem file may be gzip-compressed
(``.gz``).
"""
from __future__ import annotations

import argparse
import gzip
import json
import os
import sys

import numpy as np

from . import scenes as S

# evaluate_mbm.cpp:159-168
EVALUATE_SETTINGS = dict(num_new_configs=512, granularity=32, range=0.5, balance=2, tree_ratio=1.0,
                         dynamic_domain=1, dd_radius=4.0, dd_min_radius=1.0, dd_alpha=0.0001)
# single_mbm.cpp:128-138
SINGLE_SETTINGS = dict(num_new_configs=512, max_iters=1000000, granularity=64, range=0.5, balance=2,
                       tree_ratio=1.0, dynamic_domain=1, dd_radius=6.0, dd_min_radius=1.0, dd_alpha=0.0001)

CSV_HEADER = ("problem_name,problem_idx,solved,cost,path_length,start_tree_size,goal_tree_size,iters,wall_ns,"
              "kernel_ns,copy_ns,num_new_configs,granularity,range,balance,tree_ratio,dynamic_domain,dd_alpha,"
              "dd_radius,dd_min_radius")


def load(path: str) -> dict:
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as f:
        return json.load(f)


def problem_scene(problem: dict, family: str) -> S.Scene:
    """The obstacle tables of one MBM problem (evaluate_mbm.cpp:13-86)."""
    spheres = [S.sphere(o["position"], o["radius"]) for o in problem.get("sphere", [])]
    capsules, cuboids = [], []
    for o in problem.get("cylinder", []):
        if family == "box":
            r = float(o["radius"])
            cuboids.append(S.cuboid(o["position"], o["orientation_euler_xyz"], (r, r, r / 2.0)))
        else:
            capsules.append(S.cylinder_center(o["position"], o["orientation_euler_xyz"], o["radius"], o["length"]))
    for o in problem.get("box", []):
        cuboids.append(S.cuboid(o["position"], o["orientation_euler_xyz"], o["half_extents"]))
    return S.Scene.build(spheres=spheres, capsules=capsules, cuboids=cuboids)


def valid_problems(data: dict):
    """(family, 1-based problem index, problem dict for the planner) of every
    valid problem, in file order."""
    robot = data["robot"]
    for family, pset in data["problems"].items():
        for i, p in enumerate(pset):
            if not p.get("valid", True):
                continue
            goals = np.asarray(p["goals"], np.float32).reshape(len(p["goals"]), -1)
            yield family, p.get("index", i + 1), {"robot": robot, "start": np.asarray(p["start"], np.float32),
                                                  "goals": goals, "scene": problem_scene(p, family)}


def csv_row(family, idx, r, st) -> str:
    vals = [family, idx, int(r["solved"]), f"{r['cost']:g}", r["path_length"], r["start_tree_size"],
            r["goal_tree_size"], r["iters"], r["wall_ns"], r["kernel_ns"], r["copy_ns"], st["num_new_configs"],
            st["granularity"], f"{st['range']:g}", st["balance"], f"{st['tree_ratio']:g}", int(st["dynamic_domain"]),
            f"{st['dd_alpha']:g}", f"{st['dd_radius']:g}", f"{st['dd_min_radius']:g}"]
    return ", ".join(str(v) for v in vals)


def evaluate(data: dict, settings: dict, out_csv: str | None = None, batch: bool = False):
    """Solve every valid problem; returns [(family, idx, result, path)] and
    writes the reference's CSV when out_csv is given."""
    from . import lib
    st = dict(lib.settings_default(), **settings)
    probs = list(valid_problems(data))
    rows = []
    if batch:
        results, paths, _ = lib.solve_batch([p for _, _, p in probs], st)
        rows = [(f, i, r, pa) for (f, i, _), r, pa in zip(probs, results, paths)]
    else:
        for f, i, p in probs:
            r, pa = lib.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
            rows.append((f, i, r, pa))
    if out_csv:
        os.makedirs(os.path.dirname(os.path.abspath(out_csv)), exist_ok=True)
        with open(out_csv, "w") as f:
            f.write(CSV_HEADER + "\n")
            for fam, idx, r, _ in rows:
                f.write(csv_row(fam, idx, r, st) + "\n")
    return rows


def _quantiles(v, qs=(50, 90, 99)):
    return {f"p{q}": float(np.percentile(v, q)) for q in qs} if len(v) else {}


def summary(rows) -> dict:
    """Success rate, path cost and solve time (API wall and kernel, ms) per
    family and over all rows of `evaluate`."""
    def block(rs):
        solved = [r for r in rs if r["solved"]]
        return {"problems": len(rs), "solved": len(solved), "success_rate": len(solved) / max(1, len(rs)),
                "cost": _quantiles([r["cost"] for r in solved]),
                "wall_ms": _quantiles([r["wall_ns"] / 1e6 for r in rs]),
                "kernel_ms": _quantiles([r["kernel_ns"] / 1e6 for r in rs]),
                "iters": _quantiles([r["iters"] for r in solved])}
    fams = {}
    for fam, _, r, _ in rows:
        fams.setdefault(fam, []).append(r)
    return {"all": block([r for _, _, r, _ in rows]), "families": {f: block(rs) for f, rs in fams.items()}}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    ev = sub.add_parser("evaluate")
    ev.add_argument("robot")
    ev.add_argument("run_name")
    ev.add_argument("--problems", required=True)
    ev.add_argument("--out-dir", default="test_output")
    ev.add_argument("--batch", action="store_true")
    ev.add_argument("--max-iters", type=int, default=None)
    ev.add_argument("--max-samples", type=int, default=None)
    ev.add_argument("--summary", action="store_true")
    sg = sub.add_parser("single")
    sg.add_argument("robot")
    sg.add_argument("problem_name")
    sg.add_argument("problem_idx", type=int)
    sg.add_argument("--problems", required=True)
    args = ap.parse_args(argv)
    data = load(args.problems)
    if data["robot"] != args.robot:
        print(f"problem file is for {data['robot']}, not {args.robot}", file=sys.stderr)
        return 1
    if args.cmd == "evaluate":
        out = os.path.join(args.out_dir, f"{args.robot}_{args.run_name}.csv")
        st = dict(EVALUATE_SETTINGS)
        if args.max_iters is not None:
            st["max_iters"] = args.max_iters
        if args.max_samples is not None:
            st["max_samples"] = args.max_samples
        rows = evaluate(data, st, out, batch=args.batch)
        failed = sum(1 for _, _, r, _ in rows if not r["solved"])
        print(f"{len(rows)} problems, {failed} failed -> {out}")
        if args.summary:
            print(json.dumps(summary(rows)))
        return 0
    from . import lib
    pset = data["problems"][args.problem_name]
    p = pset[args.problem_idx - 1]
    if not p.get("valid", True):
        return 255
    st = dict(lib.settings_default(), **SINGLE_SETTINGS)
    r, path = lib.solve(args.robot, p["start"], np.asarray(p["goals"], np.float32),
                        problem_scene(p, args.problem_name), st)
    for row in path:
        print(" ".join(f"{v:g}" for v in row))
    if not r["solved"]:
        print("failed!")
    print(f"cost: {r['cost']:g}")
    print(f"time (us): {r['kernel_ns'] / 1000.0:g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
