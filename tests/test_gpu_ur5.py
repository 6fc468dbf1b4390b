"""UR5 -- a fourth robot, not in the reference, whose device code the
robot-model compiler produced from a URDF of its published kinematics and
this project's sphere file -- through the same checks as the reference's
robots: FK against the compiled table, pose and edge verdicts against the
oracle (a difference only within NEAR of contact), and plans re-validated
by the oracle, alone, batched with Panda problems in one grid, and served."""
import numpy as np
import pytest

from conftest import NEAR, edge_margins
from test_gpu_planner import _check_path, assert_same_result

pytestmark = pytest.mark.gpu


def test_ur5_fk(gpu):
    from prrtc_amd.robots import load_model
    from prrtc_amd.robots.urdf_compile import model_fk
    m = load_model("ur5")
    rng = np.random.default_rng(0)
    lo, rg = np.array(m["lower"]), np.array(m["range"])
    q = (lo + rg * rng.random((512, 6))).astype(np.float32)
    err = np.abs(gpu.fk_spheres("ur5", q) - model_fk(m, q.astype(np.float64))).max()
    assert err < 1e-5, err


@pytest.mark.parametrize("seed", [0, 1])
def test_ur5_pose_and_edge_verdicts(seed, gpu, oracle_robots):
    from prrtc_amd.robots import limits
    R = oracle_robots["ur5"]
    from prrtc_amd import workloads
    sc = workloads.ur5_scene(seed)
    lo, span = limits("ur5")
    rng = np.random.default_rng(seed + 10)
    q = (lo + span * rng.random((4096, 6), dtype=np.float32)).astype(np.float32)
    want, margin = R.check(sc, q, margin=True)
    got = gpu.check_poses("ur5", sc, q)
    assert 0.05 < want.mean() < 0.95, want.mean()
    assert np.all((got == want.astype(bool)) | (margin < NEAR))
    cand = q[want.astype(bool)][:300]
    d = rng.normal(size=cand.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    b = np.clip(cand + d * rng.uniform(0.05, 0.6, size=(len(cand), 1)), lo, lo + span).astype(np.float32)
    edges = np.stack([cand, b], 1)
    for g in (32, 128):
        we = R.validate_edges(sc, edges, g)
        ge = gpu.validate_edges("ur5", sc, edges, g)
        assert 0 < we.sum() < len(we)
        flips = np.nonzero(ge != we)[0]
        assert np.all(edge_margins(R, sc, edges[flips], g) < NEAR), flips


def _ur5_problems(n):
    from prrtc_amd import workloads
    return workloads.ur5_problems(n, seed=100)


def test_ur5_plans_alone_batched_and_mixed(gpu, oracle_robots):
    """UR5 problems solve with valid paths; in one batch with Panda problems
    (one grid for both robots) each result is bitwise its solve alone."""
    from prrtc_amd import workloads
    probs = _ur5_problems(6) + [workloads.c0_problem(i) for i in range(3)]
    st = dict(num_new_configs=64, granularity=32, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=1,
              dd_radius=4.0, dd_alpha=1e-4, dd_min_radius=1.0, max_samples=100000, max_iters=3000, seed=1)
    results, paths, _ = gpu.solve_batch(probs, st)
    solved = 0
    for p, r, path in zip(probs, results, paths):
        print(p["robot"], r["solved"], r["iters"], r["cost"], flush=True)
        r1, path1 = gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
        assert_same_result(r, path, r1, path1, p["robot"])
        if r["solved"]:
            solved += p["robot"] == "ur5"
            _check_path(oracle_robots[p["robot"]], p["scene"], path, p["start"], p["goals"], st["range"],
                        st["granularity"], r)
    assert solved >= 5, solved


def test_ur5_served(gpu, oracle_robots):
    """UR5 requests through the continuous-batching server: each result is
    bitwise its solve alone."""
    probs = _ur5_problems(4)
    st = dict(num_new_configs=64, granularity=32, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
              max_samples=100000, max_iters=2000, seed=3)
    with gpu.Server(slice_us=300) as srv:
        tickets = [srv.submit(p, st) for p in probs]
        outs = [srv.poll(t) for t in tickets]
    for p, (r, path, _) in zip(probs, outs):
        r1, path1 = gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
        assert_same_result(r, path, r1, path1, "ur5 served")
