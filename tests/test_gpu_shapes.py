"""Pass shapes never change a result.  A planner task's waves either share
a validation pass's 64 poses (splitting the obstacles) or take poses of
their own; the library picks per robot and per task.  The same problems are
solved here with the default shapes and in child processes with every
connect-chain pass allowed to go wide (PRRTC_WIDE_FROM=0, Fetch included,
whose default is never) and with steered edges kept on shared passes
(PRRTC_WIDE_EDGE=0): every result and path must be bit for bit the same."""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def solve_digests():
    """Fetch C2 problems (g 64, nnc 64: rounds leave CUs free, so chain
    passes may go wide) and Panda C1 plans (g 128: steered edges of 128
    poses), alone; one digest per solve."""
    import prrtc_amd as P
    from prrtc_amd import workloads
    P.load()
    out = []
    fst = dict(granularity=64, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0, max_samples=100000,
               num_new_configs=64, max_iters=300, seed=1)
    for p in workloads.c2_problems(6, seed=0):
        out.append(_digest(*P.solve(p["robot"], p["start"], p["goals"], p["scene"], fst)))
    pst = dict(num_new_configs=128, granularity=128, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=1,
               dd_radius=4.0, dd_alpha=1e-4, dd_min_radius=1.0, max_samples=100000, max_iters=200)
    for i, p in enumerate(workloads.c1_problems(3, seed=1)):
        out.append(_digest(*P.solve("panda", p["start"], p["goals"], p["scene"], dict(pst, seed=i))))
    return out


def _digest(r, path):
    h = hashlib.sha1()
    h.update(np.array([r["solved"], r["iters"], r["start_tree_size"], r["goal_tree_size"], r["edge_checks"]],
                      np.int64).tobytes())
    h.update(np.float32(r["cost"]).tobytes())
    h.update(path.tobytes())
    return h.hexdigest()


def _child(env_extra):
    env = dict(os.environ, **env_extra)
    code = ("import json, sys; sys.path[:0] = [%r, %r]; import test_gpu_shapes as T; "
            "print(json.dumps(T.solve_digests()))" % (ROOT, HERE))
    print(f"child process with {env_extra}", flush=True)
    res = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=100)
    assert res.returncode == 0, res.stderr[-2000:]
    return json.loads(res.stdout.strip().splitlines()[-1])


def test_pass_shapes_same_results(gpu):
    base = solve_digests()
    assert len(base) == 9
    for env_extra in ({"PRRTC_WIDE_FROM": "0"}, {"PRRTC_WIDE_EDGE": "0"}):
        got = _child(env_extra)
        bad = [i for i, (a, b) in enumerate(zip(base, got)) if a != b]
        assert not bad, f"{env_extra}: solves {bad} differ"
