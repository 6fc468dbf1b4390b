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
