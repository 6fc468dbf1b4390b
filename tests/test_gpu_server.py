"""Continuous-batching server (prrtc_server_*) on the GPU.

* every request's result (solved, iters, tree sizes, cost, path) is bitwise
  the one prrtc_solve returns for the same inputs, although the server runs
  it in short time slices among other resident problems;
* a budget that holds only a few problems at once: requests wait for
  admission, the budget is never exceeded, results are unchanged;
* error paths: footprint above the budget, invalid settings, unknown ticket.
"""
import numpy as np
import pytest

from test_gpu_planner import free_pair, point_scene

pytestmark = pytest.mark.gpu


def mixed_requests(oracle_robots, n, seed0=0):
    from prrtc_amd import scenes as S
    from prrtc_amd.robots import fixed_spheres, limits
    robots = ["panda", "fetch", "sphere", "baxter", "panda", "sphere"]
    reqs = []
    for i in range(n):
        robot = robots[i % len(robots)]
        R = oracle_robots[robot]
        if robot == "sphere":
            sc = point_scene(seed0 + i)
            lo, span = np.full(3, -5, np.float32), np.full(3, 10, np.float32)
        else:
            sc = S.scene_mixed(200 + seed0 + i, 4, 16, avoid=fixed_spheres(robot))
            lo, span = limits(robot)
        start, goal = free_pair(R, sc, lo, span, np.random.default_rng(seed0 + i))
        st = dict(num_new_configs=32 + 16 * (i % 3), granularity=32, range=0.5, balance=i % 3,
                  tree_ratio=1.0 if i % 3 == 2 else 0.5, dynamic_domain=i % 2, max_samples=40000,
                  max_iters=150, seed=seed0 + i)
        reqs.append((dict(robot=robot, start=start, goals=goal[None], scene=sc), st))
    return reqs


def assert_same(r, path, r1, path1, what):
    for k in ("solved", "iters", "start_tree_size", "goal_tree_size", "path_length", "path_nodes", "edge_checks"):
        assert r[k] == r1[k], (what, k, r[k], r1[k])
    assert np.float32(r["cost"]).tobytes() == np.float32(r1["cost"]).tobytes(), what
    assert path.tobytes() == path1.tobytes(), what


@pytest.mark.parametrize("young_rounds,crowd", [(0, False), (-1, False), (2, False), (0, True)])
def test_server_same_as_alone(gpu, oracle_robots, young_rounds, crowd):
    """Claim order (young requests first for 64 or 2 rounds, or plain
    round-robin) and crowd claims (one slot per wave while more than one
    request runs) change no result."""
    reqs = mixed_requests(oracle_robots, 12)
    alone = [gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st) for p, st in reqs]
    try:
        if crowd:
            gpu.set_tuning(None, batch_waves_per_task=4, crowd_above=1)
        with gpu.Server(slice_us=150, young_rounds=young_rounds) as srv:
            tickets = [srv.submit(p, st) for p, st in reqs]
            out = [srv.poll(t) for t in tickets]
            stats = srv.stats()
    finally:
        gpu.set_tuning(None, 0, 0, 0, 0)
    multi = 0
    for (p, _), (r1, path1), (r, path, tm), in zip(reqs, alone, out):
        print(p["robot"], r["solved"], r["iters"], tm, flush=True)
        assert_same(r, path, r1, path1, p["robot"])
        assert tm["slices"] >= 1 and tm["queued_ns"] >= 0 and tm["resident_ns"] > 0
        multi += tm["slices"] > 1
    assert multi >= 1, "no request spanned several slices; the slicing path was not exercised"
    assert stats["submitted"] == stats["completed"] == len(reqs)
    assert stats["solved"] == sum(r1["solved"] for r1, _ in alone)
    assert stats["pending"] == 0 and stats["resident"] == 0 and stats["resident_bytes"] == 0


def test_server_budget_admission(gpu, oracle_robots):
    """A budget of ~2.5 Panda problems: the rest queue, results unchanged."""
    reqs = [(p, st) for p, st in mixed_requests(oracle_robots, 12, seed0=40) if p["robot"] == "panda"]
    assert len(reqs) >= 3
    # per problem the trees dominate: 2 trees x max_samples x (7 coords + parent + radius) x 4 B
    budget = int(2.5 * 2 * 40000 * 9 * 4)
    alone = [gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st) for p, st in reqs]
    with gpu.Server(budget_bytes=budget, slice_us=200) as srv:
        tickets = [srv.submit(p, st) for p, st in reqs]
        out = [srv.poll(t) for t in tickets]
        stats = srv.stats()
    for (r1, path1), (r, path, tm) in zip(alone, out):
        assert_same(r, path, r1, path1, "panda")
    assert 0 < stats["peak_resident_bytes"] <= budget
    assert max(tm["queued_ns"] for _, _, tm in out) > 0


def test_server_errors(gpu, oracle_robots):
    (p, st), = mixed_requests(oracle_robots, 1)
    with gpu.Server(budget_bytes=1 << 20) as srv:       # 1 MiB: the trees do not fit
        with pytest.raises(gpu.PrrtcError, match="ENOMEM"):
            srv.submit(p, dict(st, max_samples=100000))
        with pytest.raises(gpu.PrrtcError, match="EINVAL"):
            srv.submit(p, dict(st, granularity=0))
        with pytest.raises(gpu.PrrtcError, match="EINVAL"):
            srv._dof[12345] = 7
            srv.poll(12345, 0)
        t = srv.submit(p, dict(st, max_samples=2000))    # fits
        r, path, tm = srv.poll(t)
        assert r["status"] == 0
