"""Real MBM problems (a subset of the reference's Panda and Fetch problem
files, tests/golden/mbm_*_subset.json) through the GPU planner.

* the reference marks every start and goal of these problems valid: the
  GPU pose checker must find them collision-free (a known answer);
* every problem is solved in one batch with the harness's granularity, and
  every returned path is re-validated edge by edge by the CPU oracle.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("robot", ["panda", "fetch"])
def test_mbm_subset_solves_with_valid_paths(robot, gpu, oracle_robots):
    from prrtc_amd import mbm
    data = mbm.load(os.path.join(GOLDEN, f"mbm_{robot}_subset.json"))
    probs = list(mbm.valid_problems(data))
    R = oracle_robots[robot]
    for fam, idx, p in probs:
        q = np.concatenate([p["start"][None], p["goals"]])
        free = gpu.check_poses(robot, p["scene"], q)
        _, margin = R.check(p["scene"], q, margin=True)
        # only a pose within 1e-6 m of contact may read differently
        assert np.all(free | (margin < 1e-6)), (fam, idx)
    st = dict(num_new_configs=256, granularity=32, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=1,
              dd_radius=4.0, dd_min_radius=1.0, dd_alpha=1e-4, max_samples=200000, max_iters=4000, seed=1)
    results, paths, stats = gpu.solve_batch([p for _, _, p in probs], st)
    solved = 0
    for (fam, idx, p), r, path in zip(probs, results, paths):
        print(robot, fam, idx, r["solved"], r["iters"], r["cost"], flush=True)
        if not r["solved"]:
            continue
        solved += 1
        assert np.array_equal(path[-1], p["start"])
        assert any(np.array_equal(path[0], gl) for gl in p["goals"])
        assert R.revalidate(p["scene"], path, 32) == -1, (fam, idx)
    assert solved >= len(probs) - 2, f"{robot}: only {solved}/{len(probs)} MBM problems solved"


@pytest.mark.parametrize("robot", ["panda", "fetch"])
def test_mbm_full_sample_batch_equals_plain(robot, gpu, oracle_robots):
    """Every 10th valid problem of the full problem file
    (tests/golden/mbm_<robot>_valid.json.gz) with the reference harness's
    settings (scripts/evaluate_mbm.cpp:154-166): every problem connects, the
    batch returns the plain solves' results bit for bit, and every path
    re-validates edge by edge on the CPU at the harness's granularity."""
    from prrtc_amd import mbm
    data = mbm.load(os.path.join(GOLDEN, f"mbm_{robot}_valid.json.gz"))
    probs = list(mbm.valid_problems(data))[::10]
    st = dict(gpu.settings_default(), **mbm.EVALUATE_SETTINGS)
    st.update(max_iters=20000, max_samples=100000)
    print(robot, len(probs), "problems", flush=True)
    results, paths, _ = gpu.solve_batch([p for _, _, p in probs], st)
    R = oracle_robots[robot]
    for (fam, idx, p), r, path in zip(probs, results, paths):
        assert r["solved"], (fam, idx)
        assert np.array_equal(path[-1], p["start"])
        assert R.revalidate(p["scene"], path, st["granularity"]) == -1, (fam, idx)
    for (fam, idx, p), r, path in list(zip(probs, results, paths))[::7]:
        r1, path1 = gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
        assert (r1["iters"], r1["cost"], r1["start_tree_size"]) == (r["iters"], r["cost"], r["start_tree_size"]), \
            (fam, idx)
        assert np.array_equal(path1, path), (fam, idx)
