"""MBM harness counterpart (prrtc_amd.mbm) on the CPU: scenes built the way
the reference's loader builds them (scripts/evaluate_mbm.cpp:13-86), the
CSV columns of evaluate_mbm.cpp:88-115."""
import os

import numpy as np

from conftest import GOLDEN


def subset(robot):
    from prrtc_amd import mbm
    return mbm.load(os.path.join(GOLDEN, f"mbm_{robot}_subset.json"))


def test_scene_tables_follow_the_reference_loader():
    from prrtc_amd import mbm
    data = subset("panda")
    for family, pset in data["problems"].items():
        for p in pset:
            sc = mbm.problem_scene(p, family)
            n_cyl, n_box = len(p["cylinder"]), len(p["box"])
            assert len(sc.spheres) == len(p["sphere"])
            if family == "box":
                # cylinders become cuboids with half extents (r, r, r/2)
                assert len(sc.capsules) == 0 and len(sc.cuboids) == n_cyl + n_box
                for o, c in zip(p["cylinder"], sc.cuboids[:n_cyl]):
                    r = o["radius"]
                    np.testing.assert_allclose(c[:3], o["position"], atol=1e-6)
                    np.testing.assert_allclose(c[12:15], [r, r, r / 2], atol=1e-7)
            else:
                assert len(sc.capsules) == n_cyl and len(sc.cuboids) == n_box
                for o, c in zip(p["cylinder"], sc.capsules):
                    p1, v = np.asarray(c[:3], np.float64), np.asarray(c[3:6], np.float64)
                    np.testing.assert_allclose(np.linalg.norm(v), o["length"], rtol=1e-5)
                    np.testing.assert_allclose(p1 + v / 2, o["position"], atol=1e-5)   # centred
                    assert abs(c[6] - o["radius"]) < 1e-7
                    np.testing.assert_allclose(c[7], 1.0 / o["length"] ** 2, rtol=1e-5)
            for o, c in zip(p["box"], sc.cuboids[n_cyl if family == "box" else 0:]):
                np.testing.assert_allclose(c[:3], o["position"], atol=1e-6)
                np.testing.assert_allclose(c[12:15], o["half_extents"], atol=1e-7)


def test_valid_problems_and_csv_format():
    from prrtc_amd import mbm
    data = subset("fetch")
    rows = list(mbm.valid_problems(data))
    assert len(rows) == 14 and all(r[2]["robot"] == "fetch" for r in rows)
    assert all(r[2]["start"].shape == (8,) and r[2]["goals"].shape[1] == 8 for r in rows)
    assert mbm.CSV_HEADER.split(",")[:3] == ["problem_name", "problem_idx", "solved"]
    assert len(mbm.CSV_HEADER.split(",")) == 20
    res = dict(solved=1, cost=2.5, path_length=3, start_tree_size=10, goal_tree_size=11, iters=4, wall_ns=5,
               kernel_ns=6, copy_ns=7)
    st = dict(mbm.EVALUATE_SETTINGS)
    line = mbm.csv_row("cage", 2, res, st)
    assert line.split(", ")[:9] == ["cage", "2", "1", "2.5", "3", "10", "11", "4", "5"]
    assert len(line.split(", ")) == 20


def test_full_fixtures_hold_every_valid_problem():
    """The compressed full-dataset fixtures hold every valid problem of the
    reference's files (699 Panda, 679 Fetch), with the subset's problems
    among them unchanged, and their starts pass the oracle's checker (the
    reference marks them valid)."""
    from oracle import oracle
    from prrtc_amd import mbm
    from prrtc_amd.robots import load_model
    oracle.build()
    for robot, n in (("panda", 699), ("fetch", 679)):
        full = mbm.load(os.path.join(GOLDEN, f"mbm_{robot}_valid.json.gz"))
        probs = list(mbm.valid_problems(full))
        assert len(probs) == n
        sub = subset(robot)
        for family, pset in sub["problems"].items():
            by_index = {p["index"]: p for p in full["problems"][family]}
            for p in pset:
                assert by_index[p["index"]] == p, (robot, family, p["index"])
        R = oracle.OracleRobot(load_model(robot))
        for fam, idx, p in probs[::25]:
            free, margin = R.check(p["scene"], p["start"][None], margin=True)
            assert free[0] or margin[0] < 1e-6, (robot, fam, idx)


def test_summary_quantiles():
    from prrtc_amd import mbm
    rows = [("a", i + 1, {"solved": i % 4 != 0, "cost": float(i), "wall_ns": 1e6 * (i + 1),
                          "kernel_ns": 5e5 * (i + 1), "iters": i}, None) for i in range(8)]
    s = mbm.summary(rows)
    assert s["all"]["problems"] == 8 and s["all"]["solved"] == 6 and s["all"]["success_rate"] == 0.75
    assert s["all"]["cost"]["p50"] == float(np.percentile([1, 2, 3, 5, 6, 7], 50))
    assert s["all"]["wall_ms"]["p50"] == 4.5 and s["families"]["a"]["solved"] == 6
