"""The HIP planner on the benchmark workload generators (prrtc_amd.workloads)
and on the reference's deterministic case:

* the hand-derived known answers of the one-block planner (tests/planner_known.py);
* C0: plain Panda problems, 12 spheres, nnc 128, g 32;
* C2: a subset of the 256 Fetch problems, batched, g 64;
* C3: a scaled Baxter / Panda / Fetch shelf mix in one launch, g 128, dd on;
* C4: a short Poisson stream of mixed requests through the server.

Every returned path is re-validated edge by edge by the oracle at the
problem's granularity; a batched or served result must be bitwise the one
prrtc_solve returns alone."""
import numpy as np
import pytest

import planner_known as K
from test_gpu_planner import _check_path, assert_same_result

pytestmark = pytest.mark.gpu


def test_known_answers_hip(gpu):
    """The one-block known answers, through prrtc_solve on the point robot."""
    def plan(scene, st):
        res, path = gpu.solve("sphere", K.START, K.GOAL[None], scene, st)
        return bool(res["solved"]), res["iters"], res["start_tree_size"], res["goal_tree_size"], path, res["cost"]
    bad = K.check_all(plan)
    assert not bad, "\n".join(bad)


def _revalidate_all(oracle_robots, problems, results, paths, settings, rng_=0.5):
    solved = 0
    for p, r, path, st in zip(problems, results, paths, settings):
        if r["solved"]:
            solved += 1
            _check_path(oracle_robots[p["robot"]], p["scene"], path, p["start"], p["goals"], st["range"],
                        st["granularity"], r)
    return solved


C0_SETTINGS = dict(num_new_configs=128, granularity=32, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
                   max_samples=100000, max_iters=2000)


def test_c0_problems(gpu, oracle_robots):
    """C0-style problems (a fresh 12-sphere scene, start and goal per seed):
    every plan re-validated; short ones compared with the oracle's
    round-synchronous planner, which must give the same path unless a pose
    within float rounding of contact flips (at most one of them)."""
    from oracle import oracle
    from prrtc_amd import workloads
    R = oracle_robots["panda"]
    solved = same = compared = 0
    for seed in range(6):
        p = workloads.c0_problem(seed)
        st = dict(C0_SETTINGS, seed=seed)
        res, path = gpu.solve("panda", p["start"], p["goals"], p["scene"], st)
        print(f"c0 seed {seed}: solved={res['solved']} rounds={res['iters']} "
              f"nodes={res['start_tree_size']}+{res['goal_tree_size']}", flush=True)
        if not res["solved"]:
            continue
        solved += 1
        _check_path(R, p["scene"], path, p["start"], p["goals"], 0.5, 32, res)
        if res["iters"] <= 4:
            oracle.set_fused(True)
            try:
                _, ores, opath = R.plan(p["scene"], p["start"], p["goals"], dict(st, max_iters=res["iters"]))
            finally:
                oracle.set_fused(False)
            compared += 1
            same += int(opath.tobytes() == path.tobytes() and ores.iters == res["iters"])
    assert solved >= 5
    assert compared >= 3 and same >= compared - 1, f"{same} of {compared} C0 plans equal the oracle's"


def test_c2_subset_batched(gpu, oracle_robots):
    """The first 16 of the 256 Fetch problems in one persistent launch:
    paths re-validated, each result bitwise the one alone."""
    from prrtc_amd import workloads
    probs = workloads.c2_problems(16, seed=0)
    st = dict(granularity=64, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0, max_samples=100000,
              max_iters=1000, num_new_configs=64, seed=1)
    results, paths, stats = gpu.solve_batch(probs, st)
    n = _revalidate_all(oracle_robots, probs, results, paths, [st] * len(probs))
    print(f"c2 subset: {n}/{len(probs)} solved, {stats['rounds']} rounds max", flush=True)
    assert n >= 12
    for p, r, path in zip(probs, results, paths):
        r1, path1 = gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
        assert_same_result(r, path, r1, path1, "c2")


def test_c3_scaled_mix(gpu, oracle_robots):
    """A 4 Baxter + 8 Panda + 4 Fetch shelf mix in one call (one persistent
    grid serves the three robots), the C3 settings: paths re-validated, each
    result bitwise the one alone."""
    from prrtc_amd import workloads
    probs = workloads.c3_problems(4, 8, 4, seed=0)
    st = dict(num_new_configs=64, granularity=128, range=0.5, balance=1, tree_ratio=0.5, dynamic_domain=1,
              dd_radius=4.0, dd_alpha=1e-4, dd_min_radius=1.0, max_samples=100000, max_iters=500, seed=1)
    results, paths, stats = gpu.solve_batch(probs, st)
    n = _revalidate_all(oracle_robots, probs, results, paths, [st] * len(probs))
    print(f"c3 mix: {n}/{len(probs)} solved, {stats['rounds']} rounds max", flush=True)
    assert n >= 10
    for p, r, path in zip(probs, results, paths):
        r1, path1 = gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
        assert_same_result(r, path, r1, path1, f"c3 {p['robot']}")


def test_c4_poisson_stream(gpu, oracle_robots):
    """A 3 s Poisson stream at 20 requests/s of serving requests (Panda :
    Fetch : Baxter = 2 : 1 : 1, 4-128 primitives): requests overlap on the
    device, each result is bitwise the one alone, every path re-validated."""
    from prrtc_amd import loadgen
    pool = loadgen.request_pool(60)
    st = dict(loadgen.SERVE_SETTINGS, max_iters=600)
    sv = loadgen.run_stream(pool, 20.0, 3.0, settings=st, slice_us=300)
    outs = sv.pop("_outs")
    print(f"c4 stream: {sv['solved']}/{sv['requests']} solved, p50 {sv['p50_ms']:.1f} ms, "
          f"p99 {sv['p99_ms']:.1f} ms, {sv['slices']} slices", flush=True)
    assert sv["rejected"] == 0 and len(outs) == sv["requests"] and sv["solved"] >= 0.7 * sv["requests"]
    for i, (rq, r, path, _) in enumerate(outs):
        rst = dict(st, seed=1 + i // len(pool))
        r1, path1 = gpu.solve(rq["robot"], rq["start"], rq["goals"], rq["scene"], rst)
        for k in ("solved", "iters", "start_tree_size", "goal_tree_size", "path_length", "edge_checks"):
            assert r[k] == r1[k], (i, k, r[k], r1[k])
        assert path.tobytes() == path1.tobytes(), i
        if r["solved"]:
            _check_path(oracle_robots[rq["robot"]], rq["scene"], path, rq["start"], rq["goals"], st["range"],
                        st["granularity"], r)
