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


def _sealed_request():
    """A point-robot problem that never connects: the start sits inside a
    sphere, so every edge from the start tree collides and the goal tree can
    never reach it (rounds run until max_iters)."""
    from prrtc_amd import scenes as S
    sc = S.Scene.build(spheres=[S.sphere((-3.5, 0, 0), 2.2)])
    p = dict(robot="sphere", start=np.array([-3.5, 0, 0], np.float32), goals=np.array([[3.5, 0.5, 0]], np.float32),
             scene=sc)
    st = dict(num_new_configs=32, granularity=16, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
              max_samples=20000, max_iters=1_000_000, seed=3)
    return p, st


def test_server_deadline(gpu, oracle_robots):
    """A resident request that never connects ends at its deadline, unsolved
    with status ETIMEDOUT (instead of running a million rounds); a request
    queued behind a full budget past its deadline is dropped without
    running; a solvable request beside them returns its solo result."""
    import time
    from prrtc_amd.lib import PRRTC_ETIMEDOUT
    sealed, sst = _sealed_request()
    (p, st), = mixed_requests(oracle_robots, 1, seed0=3)
    alone = gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
    with gpu.Server(slice_us=500, deadline_ms=60) as srv:
        t0 = time.perf_counter()
        ts = srv.submit(sealed, sst)
        tp = srv.submit(p, st)
        r, path, tm = srv.poll(ts)
        waited = time.perf_counter() - t0
        r1, path1, _ = srv.poll(tp)
        stats = srv.stats()
    print(f"sealed request: status {r['status']}, {tm['slices']} slices, {waited * 1e3:.1f} ms", flush=True)
    assert not r["solved"] and r["status"] == PRRTC_ETIMEDOUT and r["start_tree_size"] == 1
    assert 0.06 <= waited < 2.0
    assert_same(r1, path1, *alone, "beside a timed-out request")
    assert stats["timed_out"] == 1 and stats["completed"] == 2
    # a budget that holds one sealed problem: the second waits in the queue;
    # whether it is dropped there or admitted once the first has stopped,
    # neither outlives its deadline by more than a few slices
    one = int(1.5 * 2 * 20000 * 5 * 4)
    with gpu.Server(slice_us=500, deadline_ms=40, budget_bytes=one) as srv:
        ta, tb = srv.submit(sealed, sst), srv.submit(sealed, sst)
        ra, _, tma = srv.poll(ta)
        rb, _, tmb = srv.poll(tb)
        stats = srv.stats()
    assert ra["status"] == rb["status"] == PRRTC_ETIMEDOUT and stats["timed_out"] == 2
    for tm in (tma, tmb):
        assert 40e6 <= tm["queued_ns"] + tm["resident_ns"] <= 40e6 + 30e6, tm
    assert tmb["queued_ns"] > 0 and stats["peak_resident_bytes"] <= one


def test_server_cancel_queued(gpu):
    """A request queued behind a full budget and cancelled there ends at
    once, never admitted (no slice, no resident time), status ECANCELED."""
    from prrtc_amd.lib import PRRTC_ECANCELED
    sealed, sst = _sealed_request()
    one = int(1.5 * 2 * 20000 * 5 * 4)
    with gpu.Server(slice_us=500, budget_bytes=one) as srv:
        ta, tb = srv.submit(sealed, sst), srv.submit(sealed, sst)
        srv.cancel(tb)
        rb, _, tmb = srv.poll(tb)
        srv.cancel(ta)
        ra, _, _ = srv.poll(ta)
    assert rb["status"] == ra["status"] == PRRTC_ECANCELED
    assert tmb["slices"] == 0 and tmb["resident_ns"] == 0


def test_server_cancel(gpu, oracle_robots):
    """prrtc_server_cancel ends a resident request (status ECANCELED) within
    a few slices; cancelling a finished request keeps its result; an unknown
    ticket is EINVAL."""
    import time
    from prrtc_amd.lib import PRRTC_ECANCELED
    sealed, sst = _sealed_request()
    (p, st), = mixed_requests(oracle_robots, 1, seed0=4)
    alone = gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
    with gpu.Server(slice_us=500) as srv:
        ts = srv.submit(sealed, sst)
        tp = srv.submit(p, st)
        time.sleep(0.03)
        assert srv.poll(ts, 0) is None                  # still running
        t0 = time.perf_counter()
        srv.cancel(ts)
        r, _, tm = srv.poll(ts)
        took = time.perf_counter() - t0
        r1, path1, _ = srv.poll(tp)
        stats = srv.stats()
        with pytest.raises(gpu.PrrtcError, match="EINVAL"):
            srv.cancel(987654)
    print(f"cancelled after {tm['slices']} slices, {took * 1e3:.1f} ms after the call", flush=True)
    assert not r["solved"] and r["status"] == PRRTC_ECANCELED and took < 1.0
    assert_same(r1, path1, *alone, "beside a cancelled request")
    assert stats["cancelled"] == 1
    with gpu.Server(slice_us=500) as srv:               # cancel after the result: it stands
        t = srv.submit(p, st)
        while srv.stats()["completed"] < 1:
            time.sleep(0.005)
        srv.cancel(t)
        r2, path2, _ = srv.poll(t)
    assert_same(r2, path2, *alone, "cancelled after finishing")
