"""prrtc_solve_noalloc (the allocation-free prrtc_solve, device state in the
library's static scratch array) against prrtc_solve, through the C ABI:
bitwise the same results on two calls in a row, the second with a smaller
problem (fewer obstacles, fewer samples), so what the first call left in the
array must not leak into the second; and a problem too big for the array
is refused with ENOMEM."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("solved", "iters", "start_tree_size", "goal_tree_size", "path_length", "path_nodes", "edge_checks",
          "pose_checks", "status")


def _same(a, pa, b, pb, what):
    for k in FIELDS:
        assert a[k] == b[k], (what, k, a[k], b[k])
    assert np.float32(a["cost"]).tobytes() == np.float32(b["cost"]).tobytes(), what
    assert pa.tobytes() == pb.tobytes(), what


def test_noalloc_equals_solve(gpu):
    from prrtc_amd import scenes as S
    from prrtc_amd import workloads
    big = workloads.c2_problems(1, seed=3)[0]                        # Fetch, 8-64 obstacles
    st_big = dict(num_new_configs=64, granularity=64, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=1,
                  dd_radius=4.0, dd_alpha=1e-4, dd_min_radius=1.0, max_samples=60000, max_iters=1500, seed=5)
    small_sc = S.Scene.build(spheres=[S.sphere((0.0, 0.0, 0.0), 1.2)])
    small = dict(robot="sphere", start=np.array([-3.0, 0.3, 0.1], np.float32),
                 goals=np.array([[3.0, -0.2, 0.4]], np.float32), scene=small_sc)
    st_small = dict(num_new_configs=16, granularity=16, range=0.5, balance=1, tree_ratio=0.5, dynamic_domain=0,
                    max_samples=4000, max_iters=300, seed=2)
    for p, st in ((big, st_big), (small, st_small), (big, st_big)):
        want, wpath = gpu.solve(p["robot"], p["start"], p["goals"], p["scene"], st)
        got, gpath = gpu.solve_noalloc(p["robot"], p["start"], p["goals"], p["scene"], st)
        print(p["robot"], got["solved"], got["iters"], got["start_tree_size"], got["goal_tree_size"], flush=True)
        _same(got, gpath, want, wpath, p["robot"])
    assert want["solved"]


def test_noalloc_rejects_oversized(gpu):
    from prrtc_amd import scenes as S
    sc = S.Scene.build(spheres=[S.sphere((0.0, 0.0, 0.0), 1.0)])
    st = dict(num_new_configs=16, granularity=16, range=0.5, max_samples=5_000_000, max_iters=10)
    with pytest.raises(gpu.PrrtcError, match="ENOMEM"):
        gpu.solve_noalloc("panda", np.zeros(7, np.float32), np.zeros((1, 7), np.float32), sc, st)
    with pytest.raises(gpu.PrrtcError, match="static scratch"):
        gpu.solve_noalloc("panda", np.zeros(7, np.float32), np.zeros((1, 7), np.float32), sc, st)
