"""GPU planner (prrtc_solve / prrtc_solve_batch) against the CPU oracle.

* point robot: every quantity bit for bit equal to the oracle's
  round-synchronous planner (oracle in the kernels' fmaf rounding);
* Panda / Fetch / Baxter: every returned path re-validated edge by edge by
  the oracle at the same granularity, endpoints and cost checked;
* a problem gives the same result alone and inside a batch of others;
* forced failures (iteration budget, node budget, sealed start).
"""
import numpy as np
import pytest

from conftest import ROBOTS

pytestmark = pytest.mark.gpu


def point_scene(seed):
    from prrtc_amd import scenes as S
    rng = np.random.default_rng(seed)
    return S.Scene.build(
        spheres=[S.sphere(rng.uniform(-3, 3, 3), rng.uniform(0.4, 1.2)) for _ in range(5)],
        cuboids=[S.cuboid(rng.uniform(-3, 3, 3), rng.uniform(-1, 1, 3), rng.uniform(0.2, 0.9, 3)) for _ in range(3)],
        capsules=[S.capsule_endpoints(rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 3), 0.3) for _ in range(2)],
        z_capsules=[S.capsule_endpoints((1.5, -1.5, -2.0), (1.5, -1.5, 2.0), 0.4)],
        z_cuboids=[S.cuboid((-1.5, 1.5, 0.0), (0.0, 0.0, 0.4), (0.5, 0.3, 1.0))])


def free_pair(R, scene, lo, span, rng, far=None, tries=16):
    """Two collision-free poses (the farthest pair of a draw when `far`), or
    None when the scene leaves (almost) no free pose."""
    for _ in range(tries):
        q = (lo + span * rng.random((256, len(lo)), dtype=np.float32)).astype(np.float32)
        q = q[R.check(scene, q)]
        if len(q) >= 2:
            if far is None:
                return q[0], q[1]
            d = np.linalg.norm(q[:, None] - q[None], axis=-1)
            i, j = np.unravel_index(np.argmax(d), d.shape)
            return q[i], q[j]
    return None


POINT_SETTINGS = [
    dict(num_new_configs=32, granularity=16, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0),
    dict(num_new_configs=64, granularity=64, range=0.7, balance=1, tree_ratio=0.5, dynamic_domain=1,
         dd_alpha=1e-2, dd_radius=2.0, dd_min_radius=0.5),
    dict(num_new_configs=100, granularity=7, range=0.4, balance=0, tree_ratio=0.5, dynamic_domain=1),
    dict(num_new_configs=16, granularity=128, range=1.0, balance=2, tree_ratio=1.0, dynamic_domain=0),
]


@pytest.mark.parametrize("case", range(len(POINT_SETTINGS)))
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_point_robot_bit_exact(case, seed, gpu, oracle_robots):
    from oracle import oracle
    R = oracle_robots["sphere"]
    sc = point_scene(seed)
    rng = np.random.default_rng(100 + seed)
    lo, span = np.full(3, -5, np.float32), np.full(3, 10, np.float32)
    start, goal = free_pair(R, sc, lo, span, rng, far=True)
    goals = np.stack([goal, free_pair(R, sc, lo, span, rng)[0]]) if case == 1 else goal[None]
    st = dict(POINT_SETTINGS[case], max_samples=50000, max_iters=400, seed=seed + 7)
    oracle.set_fused(True)
    try:
        rc, ores, opath = R.plan(sc, start, goals, st, mode=0)
    finally:
        oracle.set_fused(False)
    res, path = gpu.solve("sphere", start, goals, sc, st)
    assert rc == 0
    assert res["solved"] == ores.solved
    assert res["iters"] == ores.iters
    assert (res["start_tree_size"], res["goal_tree_size"]) == (ores.start_tree_size, ores.goal_tree_size)
    assert res["path_length"] == ores.path_length
    assert np.float32(res["cost"]).tobytes() == np.float32(ores.cost).tobytes()
    assert path.tobytes() == opath.tobytes()


def test_long_connection_trace(gpu, oracle_robots):
    """A connection whose trees hold long walks (range 0.1 across 15.6 units
    around three large spheres: a 251-row path), so the commit's trace runs
    past the 64 rows it keeps in registers: rows, cost and counts bit for bit
    the oracle's."""
    from oracle import oracle
    from prrtc_amd import scenes as S
    R = oracle_robots["sphere"]
    sc = S.Scene.build(spheres=[S.sphere((0, 0, 0), 3.0), S.sphere((2.5, -2.5, 2.0), 1.5),
                                S.sphere((-2.5, 2.5, -2.0), 1.5)])
    start = np.array([-4.5, -4.5, -4.5], np.float32)
    goal = np.array([4.5, 4.5, 4.5], np.float32)
    st = dict(num_new_configs=8, granularity=4, range=0.1, balance=2, tree_ratio=1.0, dynamic_domain=0,
              max_samples=50000, max_iters=400, seed=3)
    oracle.set_fused(True)
    try:
        rc, ores, opath = R.plan(sc, start, goal[None], st, mode=0)
    finally:
        oracle.set_fused(False)
    assert rc == 0 and ores.solved and len(opath) > 2 * 64 + 2
    res, path = gpu.solve("sphere", start, goal[None], sc, st)
    assert res["solved"] and res["iters"] == ores.iters
    assert (res["start_tree_size"], res["goal_tree_size"]) == (ores.start_tree_size, ores.goal_tree_size)
    assert res["path_length"] == ores.path_length
    assert np.float32(res["cost"]).tobytes() == np.float32(ores.cost).tobytes()
    assert path.tobytes() == opath.tobytes()


def _check_path(R, scene, path, start, goals, rng_, g, res):
    assert np.array_equal(path[-1], start)
    assert any(np.array_equal(path[0], gl) for gl in goals)
    seg = np.linalg.norm(np.diff(path.astype(np.float64), axis=0), axis=1)
    assert np.all(seg <= rng_ * (1 + 1e-4) + 1e-5)
    assert abs(seg.sum() - res["cost"]) <= 1e-3 * max(1.0, res["cost"])
    assert res["path_length"] == len(path) - 2
    assert R.revalidate(scene, path, g) == -1, "path has a colliding edge"


@pytest.mark.parametrize("robot", ROBOTS)
def test_paths_revalidate(robot, gpu, oracle_robots):
    from prrtc_amd import scenes as S
    from prrtc_amd.robots import fixed_spheres, limits
    R = oracle_robots[robot]
    lo, span = limits(robot)
    solved = 0
    same = 0
    compared = 0
    tried = 0
    for seed in range(16):
        if tried == 4:
            break
        sc = S.scene_mixed(seed, 8, 24, avoid=fixed_spheres(robot))
        rng = np.random.default_rng(seed)
        pair = free_pair(R, sc, lo, span, rng)
        if pair is None:
            continue
        tried += 1
        start, goal = pair
        st = dict(num_new_configs=64, granularity=32, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
                  max_samples=50000, max_iters=200, seed=seed)
        res, path = gpu.solve(robot, start, goal, sc, st)
        print(f"{robot} seed {seed}: solved={res['solved']} rounds={res['iters']} nodes="
              f"{res['start_tree_size']}+{res['goal_tree_size']}", flush=True)
        if res["solved"]:
            solved += 1
            _check_path(R, sc, path, start, goal[None], 0.5, 32, res)
            if res["iters"] <= 8:
                # the oracle's round-synchronous planner, same rounds: the path is
                # the same unless a pose within float rounding of contact flips
                rc, ores, opath = R.plan(sc, start, goal, dict(st, max_iters=res["iters"]), mode=0)
                same += int(opath.tobytes() == path.tobytes())
                compared += 1
    assert solved >= 3
    assert same >= compared - 1, f"only {same} of {compared} paths match the oracle's planner"
    print(f"{robot}: {solved}/4 solved, {same} paths identical to the oracle's round-synchronous planner")


def mixed_batch(oracle_robots):
    from prrtc_amd import scenes as S
    from prrtc_amd.robots import fixed_spheres, limits
    probs, sts = [], []
    for i, robot in enumerate(["panda", "fetch", "panda", "baxter", "sphere", "fetch"]):
        R = oracle_robots[robot]
        if robot == "sphere":
            sc = point_scene(i)
            lo, span = np.full(3, -5, np.float32), np.full(3, 10, np.float32)
        else:
            sc = S.scene_mixed(50 + i, 4, 16, avoid=fixed_spheres(robot))
            lo, span = limits(robot)
        start, goal = free_pair(R, sc, lo, span, np.random.default_rng(i))
        probs.append(dict(robot=robot, start=start, goals=goal[None], scene=sc))
        sts.append(dict(num_new_configs=32 + 16 * (i % 3), granularity=32, range=0.5, balance=i % 3,
                        tree_ratio=1.0 if i % 3 == 2 else 0.5, dynamic_domain=i % 2, max_samples=40000,
                        max_iters=150, seed=i))
    return probs, sts


def assert_same_result(r, path, r1, path1, what):
    for k in ("solved", "iters", "start_tree_size", "goal_tree_size", "path_length", "edge_checks"):
        assert r[k] == r1[k], (what, k, r[k], r1[k])
    assert np.float32(r["cost"]) == np.float32(r1["cost"]), what
    assert path.tobytes() == path1.tobytes(), what


def test_batch_same_as_alone(gpu, oracle_robots):
    """Each problem's result is bitwise the same alone and inside a mixed batch."""
    probs, sts = mixed_batch(oracle_robots)
    results, paths, stats = gpu.solve_batch(probs, sts)
    assert stats["solved"] == sum(r["solved"] for r in results)
    for p, st, r, path in zip(probs, sts, results, paths):
        r1, path1 = gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
        assert_same_result(r, path, r1, path1, p["robot"])


def test_graph_rounds_same_as_persistent(gpu, oracle_robots):
    """prrtc_solve_batch_rounds (one task grid and one commit grid per round,
    replayed from a captured graph) returns bitwise the persistent batch's
    results, with several graph replays (problems run up to 150 rounds)."""
    probs, sts = mixed_batch(oracle_robots)
    results, paths, stats = gpu.solve_batch(probs, sts)
    gres, gpaths, gstats = gpu.solve_batch(probs, sts, graph_rounds=True)
    assert gstats["solved"] == stats["solved"] and gstats["rounds"] == stats["rounds"]
    assert stats["rounds"] > 8, "every problem finished within one graph replay"
    for p, r, path, r1, path1 in zip(probs, gres, gpaths, results, paths):
        assert_same_result(r, path, r1, path1, p["robot"])


def test_forced_failures(gpu, oracle_robots):
    from prrtc_amd import scenes as S
    sc = S.Scene.build(spheres=[S.sphere((0, 0, 0), 1.5)])
    start = np.array([-3.5, 0, 0], np.float32)
    goal = np.array([3.5, 0.5, 0], np.float32)
    base = dict(num_new_configs=8, granularity=8, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
                max_samples=20000, max_iters=2, seed=3)
    res, path = gpu.solve("sphere", start, goal, sc, base)                  # iteration budget
    assert not res["solved"] and res["iters"] == 0 and len(path) == 0
    sealed = S.Scene.build(spheres=[S.sphere((-3.5, 0, 0), 2.2)])
    res, _ = gpu.solve("sphere", start, goal, sealed, dict(base, max_iters=50))   # every edge collides
    assert not res["solved"] and res["start_tree_size"] == 1
    res, _ = gpu.solve("sphere", start, goal, sc, dict(base, max_iters=1000, max_samples=12))  # node budget
    assert not res["solved"]
    with pytest.raises(gpu.PrrtcError):
        gpu.solve("sphere", start, goal, sc, dict(base, granularity=0))
    with pytest.raises(gpu.PrrtcError):
        gpu.solve("sphere", np.array([np.nan, 0, 0], np.float32), goal, sc, base)


@pytest.mark.parametrize("waves", [2, 4])
def test_task_waves_same_results(waves, gpu, oracle_robots):
    """A task spread over several waves (shared nearest-node scans, obstacle
    tables split between the waves of each validation pass) returns bitwise
    the single-wave result, alone and inside a one-wave batch; so does a
    batch of `waves`-wave workgroups with crowd claims forced on (each wave
    runs its own slot while more than one problem runs, the waves share a
    slot for the last problem)."""
    from prrtc_amd import scenes as S
    from prrtc_amd.robots import fixed_spheres, limits
    cases = []
    # 9 more point-robot problems make the point-robot group a real batch
    for i in range(21):
        robot = ["sphere", "panda", "fetch", "baxter"][i % 4] if i < 12 else "sphere"
        R = oracle_robots[robot]
        if robot == "sphere":
            sc = point_scene(10 + i)
            lo, span = np.full(3, -5, np.float32), np.full(3, 10, np.float32)
        else:
            sc = S.scene_mixed(300 + i, 8, 24, avoid=fixed_spheres(robot))
            lo, span = limits(robot)
        start, goal = free_pair(R, sc, lo, span, np.random.default_rng(70 + i), far=True)
        st = dict(num_new_configs=48, granularity=[16, 64, 128, 100][i % 4], range=0.5, balance=i % 3,
                  tree_ratio=0.5, dynamic_domain=i % 2, max_samples=30000, max_iters=120, seed=i)
        cases.append((robot, start, goal[None], sc, st))
    try:
        gpu.set_tuning(None, waves_per_task=1, batch_waves_per_task=1)
        ref = [gpu.solve(*c) for c in cases]
        gpu.set_tuning(None, waves_per_task=waves, batch_waves_per_task=1)
        assert gpu.get_tuning("panda")["waves_per_task"] == waves
        got = [gpu.solve(*c) for c in cases]
        problems = [dict(robot=c[0], start=c[1], goals=c[2], scene=c[3]) for c in cases]
        batch, bpaths, _ = gpu.solve_batch(problems, [c[4] for c in cases])
        gpu.set_tuning(None, waves_per_task=waves, batch_waves_per_task=waves, crowd_above=1)
        assert gpu.get_tuning("fetch")["crowd_above"] == 1
        crowd, cpaths, _ = gpu.solve_batch(problems, [c[4] for c in cases])
    finally:
        gpu.set_tuning(None, 0, 0, 0, 0)
    for (robot, *_), (r1, p1), (r2, p2), rb, pb, rc, pc in zip(cases, ref, got, batch, bpaths, crowd, cpaths):
        print(robot, r1["solved"], r1["iters"], r1["start_tree_size"], r1["goal_tree_size"], flush=True)
        for k in ("solved", "iters", "start_tree_size", "goal_tree_size", "path_length", "edge_checks"):
            assert r1[k] == r2[k] == rb[k] == rc[k], (robot, k, r1[k], r2[k], rb[k], rc[k])
        assert p1.tobytes() == p2.tobytes() == pb.tobytes() == pc.tobytes(), robot
