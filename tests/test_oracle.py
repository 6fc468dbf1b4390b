"""Pin the CPU oracle: golden vectors from the reference's own FK/checker,
closed forms, brute force, and planted mistakes that must be caught."""
import copy
import math
from fractions import Fraction

import numpy as np
import pytest

from conftest import ROBOTS, golden, golden_scene

NEAR = 1e-6   # metres: poses this close to contact may flip verdict


@pytest.mark.parametrize("robot", ROBOTS)
def test_fk_matches_reference(robot, oracle_robots):
    g = golden(f"fk_{robot}.npz")
    c = oracle_robots[robot].fk(g["q"])
    err = np.abs(c - g["centres"]).max()
    assert err < 1e-5, f"{robot}: sphere centres off by {err} m"
    # the radii of the model tables are the reference's literals
    from prrtc_amd.robots import load_model
    r = np.array([s["r"] for s in load_model(robot)["spheres"]], dtype=np.float32)
    np.testing.assert_array_equal(r, g["radii"])


@pytest.mark.parametrize("robot", ROBOTS)
@pytest.mark.parametrize("scene", ["spheres12", "obb_cyl", "shelf", "ztables"])
def test_pose_verdicts_match_reference(robot, scene, oracle_robots):
    g = golden(f"verdict_{robot}.npz")
    sc = golden_scene(g, scene)
    free, margin = oracle_robots[robot].check(sc, g["poses"], margin=True)
    ref = g[f"{scene}__free"]
    mism = free != ref
    near = margin < NEAR
    assert not np.any(mism & ~near), f"{(mism & ~near).sum()} verdicts differ away from contact"
    # report near-contact flips (expected: none or very few)
    assert mism.sum() <= max(2, int(0.001 * len(ref)))


@pytest.mark.parametrize("robot", ROBOTS)
@pytest.mark.parametrize("g", [32, 64, 128])
def test_edge_verdicts_match_reference(robot, g, oracle_robots):
    e = golden(f"edge_{robot}.npz")
    sc = golden_scene(e, "scene")
    free = oracle_robots[robot].validate_edges(sc, e["edges"], g)
    assert (free != e[f"free_g{g}"]).sum() == 0


# ---------------------------------------------------------------- closed forms

def _point_obstacle_distance(kind, row, p):
    from prrtc_amd.scenes import obstacle_distance
    return obstacle_distance(kind, row, p)


@pytest.mark.parametrize("kind", ["sphere", "capsule", "z_capsule", "cuboid", "z_cuboid"])
def test_point_robot_closed_form(kind, oracle_robots):
    """The point robot's check is one env test (utils.cuh:330-334): a radius-1
    sphere at q collides iff the textbook point-obstacle distance is < 1."""
    from prrtc_amd import scenes as S
    rng = np.random.default_rng(7)
    if kind == "sphere":
        row = S.sphere((0.3, -0.2, 0.5), 0.7)
    elif kind == "capsule":
        row = S.capsule_endpoints((-0.5, 0.2, 0.1), (0.8, -0.4, 0.9), 0.3)
    elif kind == "z_capsule":
        row = S.capsule_endpoints((0.1, 0.2, -0.5), (0.1, 0.2, 1.0), 0.3)
    elif kind == "cuboid":
        row = S.cuboid((0.2, 0.1, -0.3), (0.3, -0.5, 1.1), (0.4, 0.2, 0.6))
    else:
        row = S.cuboid((0.2, 0.1, -0.3), (0.0, 0.0, 0.7), (0.4, 0.2, 0.6))
    table = {"sphere": "spheres", "capsule": "capsules", "z_capsule": "z_capsules", "cuboid": "cuboids",
             "z_cuboid": "z_cuboids"}[kind]
    sc = S.Scene.build(**{table: [row]})
    q = rng.uniform(-2.5, 2.5, size=(4000, 3)).astype(np.float32)
    free, margin = oracle_robots["sphere"].check(sc, q, margin=True)
    d = np.array([_point_obstacle_distance(kind, row, p) for p in q])
    truth_free = d >= 1.0
    mism = free != truth_free
    assert not np.any(mism & (np.abs(d - 1.0) > 1e-5))
    assert 0.05 < free.mean() < 0.95


def test_halton_slot0_is_radical_inverse():
    """Slot 0 uses the unpermuted bases 3, 5, 7, ... (pRRTC.cu:61-65); the k-th
    draw is the radical inverse phi_b(k), an exact fraction n/d, and the
    float algorithm returns it correctly rounded."""
    from oracle.oracle import halton
    primes = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]
    n = 700
    out = halton(14, seed=1, slot=0, n=n)

    def phi(b, k):
        f, inv = Fraction(0), Fraction(1, b)
        while k:
            f += (k % b) * inv
            k //= b
            inv /= b
        return f

    for d, b in enumerate(primes):
        want = np.array([np.float32(float(phi(b, k))) for k in range(1, n + 1)], dtype=np.float32)
        np.testing.assert_array_equal(out[:, d], want)


def test_halton_other_slots_permute_bases():
    from oracle.oracle import halton
    first = halton(14, seed=9, slot=5, n=1)[0]
    # phi_b(1) = 1/b: the first draw reveals each dimension's base
    bases = sorted(round(1.0 / v) for v in first)
    assert set(bases) <= {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59}
    assert len(set(bases)) == 14
    a = halton(7, seed=9, slot=5, n=50)
    b = halton(7, seed=9, slot=5, n=50)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(halton(7, seed=9, slot=5, n=5), halton(7, seed=9, slot=6, n=5))


def test_factory_closed_forms():
    from prrtc_amd import scenes as S
    rho, theta, phi = 0.3, -0.7, 1.9
    R = S.euler_xyz_matrix(rho, theta, phi).astype(np.float64)
    cx, sx = math.cos(rho), math.sin(rho)
    cy, sy = math.cos(theta), math.sin(theta)
    cz, sz = math.cos(phi), math.sin(phi)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    np.testing.assert_allclose(R, Rz @ Ry @ Rx, atol=1e-6)
    cap = S.cylinder_center((0.1, 0.2, 0.3), (rho, theta, phi), 0.05, 0.4)
    p1, v = cap[:3].astype(np.float64), cap[3:6].astype(np.float64)
    np.testing.assert_allclose(p1, np.array([0.1, 0.2, 0.3]) + R[:, 2] * 0.2, atol=1e-6)
    np.testing.assert_allclose(p1 + v, np.array([0.1, 0.2, 0.3]) - R[:, 2] * 0.2, atol=1e-6)
    assert abs(cap[7] - 1.0 / np.dot(v, v)) < 1e-5 * cap[7]
    box = S.cuboid((0, 0, 0), (rho, theta, phi), (1, 2, 3))
    np.testing.assert_allclose(box[3:12].reshape(3, 3), R.T, atol=1e-6)


# ------------------------------------------------------- planted mistakes

@pytest.mark.parametrize("mistake", ["offset_sign", "radius", "swap_axis"])
def test_planted_mistakes_are_caught(mistake):
    """A plausible slip in the model tables must fail the golden checks."""
    from oracle.oracle import OracleRobot
    from prrtc_amd.robots import load_model
    m = copy.deepcopy(load_model("panda"))
    if mistake == "offset_sign":
        m["spheres"][30]["local"][1] = -m["spheres"][30]["local"][1]
    elif mistake == "radius":
        for s in m["spheres"]:
            s["r"] *= 1.05
    else:
        rot = np.array(m["joints"][3]["rot"])
        m["joints"][3]["rot"] = rot.T.tolist()
    R = OracleRobot(m)
    g = golden("fk_panda.npz")
    fk_err = np.abs(R.fk(g["q"]) - g["centres"]).max()
    v = golden("verdict_panda.npz")
    flips = 0
    for name in ("spheres12", "obb_cyl", "shelf", "ztables"):
        flips += int((R.check(golden_scene(v, name), v["poses"]) != v[f"{name}__free"]).sum())
    assert fk_err > 1e-5 or flips > 0


# ------------------------------------------------------------ planner oracle

def _path_invariants(R, scene, path, start, goal, rng_, g, cost):
    assert np.array_equal(path[0], goal)
    assert np.array_equal(path[-1], start)
    seg = np.linalg.norm(np.diff(path.astype(np.float64), axis=0), axis=1)
    assert np.all(seg <= rng_ * (1 + 1e-4) + 1e-6)
    assert abs(seg.sum() - cost) <= 1e-3 * max(1.0, cost)
    assert R.revalidate(scene, path, g) == -1


@pytest.mark.parametrize("mode", [0, 1])
def test_planner_oracle_point_robot(mode, oracle_robots):
    from prrtc_amd import scenes as S
    R = oracle_robots["sphere"]
    sc = S.Scene.build(spheres=[S.sphere((0, 0, 0), 1.5)], cuboids=[S.cuboid((0, 2.5, 0), (0, 0, 0), (2, 0.3, 2))])
    start = np.array([-3.5, 0, 0], np.float32)
    goal = np.array([3.5, 0.5, 0], np.float32)
    st = dict(num_new_configs=16, granularity=16, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
              max_samples=20000, max_iters=2000, seed=3)
    rc, res, path = R.plan(sc, start, goal, st, mode=mode)
    assert rc == 0 and res.solved
    assert res.path_length == len(path) - 2
    _path_invariants(R, sc, path, start, goal, 0.5, 16, res.cost)


@pytest.mark.parametrize("mode", [0, 1])
def test_planner_oracle_panda(mode, oracle_robots):
    from prrtc_amd import scenes as S
    from prrtc_amd.robots import limits
    R = oracle_robots["panda"]
    sc = S.scene_spheres(0)
    lo, span = limits("panda")
    rng = np.random.default_rng(0)
    free = []
    while len(free) < 2:
        q = (lo + span * rng.random((64, 7), dtype=np.float32)).astype(np.float32)
        free += list(q[R.check(sc, q)])
    start, goal = free[0], free[1]
    st = dict(num_new_configs=32, granularity=32, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
              max_samples=20000, max_iters=300, seed=1)
    rc, res, path = R.plan(sc, start, goal, st, mode=mode)
    assert rc == 0 and res.solved
    _path_invariants(R, sc, path, start, goal, 0.5, 32, res.cost)


def test_rrt_connect_baseline(oracle_robots):
    from prrtc_amd import scenes as S
    R = oracle_robots["sphere"]
    sc = S.Scene.build(spheres=[S.sphere((0, 0, 0), 1.5)])
    start = np.array([-3.5, 0, 0], np.float32)
    goal = np.array([3.5, 0.5, 0], np.float32)
    rc, res, path = R.rrt_connect(sc, start, goal, 16, 0.5, 5000, 11)
    assert rc == 0 and res.solved
    assert np.array_equal(path[0], goal) and np.array_equal(path[-1], start)
    assert R.revalidate(sc, path, 16) == -1


def test_failure_paths(oracle_robots):
    """Forced failures: a tiny iteration budget, a start sealed inside an
    obstacle (every edge fails), and a tiny sample budget."""
    from prrtc_amd import scenes as S
    R = oracle_robots["sphere"]
    start = np.array([-3.5, 0, 0], np.float32)
    goal = np.array([3.5, 0.5, 0], np.float32)
    base = dict(num_new_configs=8, granularity=8, range=0.5, balance=2, tree_ratio=1.0, dynamic_domain=0,
                max_samples=20000, max_iters=2, seed=3)
    sc = S.Scene.build(spheres=[S.sphere((0, 0, 0), 1.5)])
    rc, res, _ = R.plan(sc, start, goal, base)
    assert rc == 0 and not res.solved and res.iters == 0
    sealed = S.Scene.build(spheres=[S.sphere((-3.5, 0, 0), 2.2)])   # start pose inside: every edge collides
    rc, res, _ = R.plan(sealed, start, goal, dict(base, max_iters=50))
    assert rc == 0 and not res.solved and res.start_tree_size == 1
    rc, res, _ = R.plan(sc, start, goal, dict(base, max_iters=1000, max_samples=12))
    assert rc == 0 and not res.solved
