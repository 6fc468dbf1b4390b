"""Synthetic planning workloads of the benchmark configurations.

    c0_problem   Panda, 12 spheres, one problem                         (plain)
    c1_scene     Panda, 24 oriented cuboids + 8 vertical cylinders      (microbench / plan)
    c1_edges     1M random edges, endpoints in the limits, length <= 0.5 rad
    c2_problems  256 Fetch problems, 8-64 mixed obstacles each          (batched)
    c3_problems  128 Baxter + 256 Panda + 128 Fetch, shelves + spheres   (uneven mix)
    ur5_problems UR5 (the compiled fourth robot), 12 mixed obstacles in reach
    serving_request  one request of the Poisson serving stream

Scenes come from prrtc_amd.scenes (SplitMix64-seeded).  Obstacles that
touch a robot's fixed base spheres are redrawn, and a scene in which almost
no pose is collision-free is redrawn with the next seed.  Starts and goals
are uniform in the joint limits, resampled until the HIP pose check says they
are collision-free.
"""
from __future__ import annotations

import numpy as np

from . import lib
from . import scenes as S
from .robots import fixed_spheres, limits

POINT_LIMITS = (np.full(3, -5.0, np.float32), np.full(3, 10.0, np.float32))


def robot_limits(robot):
    return POINT_LIMITS if robot == "sphere" else limits(robot)


def free_poses(robot, scene, n, rng, max_draws=64, batch=4096):
    """n collision-free poses uniform in the limits (None if the scene is
    too cluttered to find them)."""
    lo, span = robot_limits(robot)
    out = []
    for _ in range(max_draws):
        q = (lo + span * rng.random((batch, len(lo)), dtype=np.float32)).astype(np.float32)
        out.extend(q[lib.check_poses(robot, scene, q)])
        if len(out) >= n:
            return np.stack(out[:n])
    return None


def feasible_scene(robot, make, seed, rng, min_free=0.005):
    """Draw scenes make(seed), make(seed + 100003), ... until at least
    `min_free` of uniform poses are collision-free."""
    lo, span = robot_limits(robot)
    for k in range(64):
        sc = make(seed + 100003 * k)
        q = (lo + span * rng.random((2048, len(lo)), dtype=np.float32)).astype(np.float32)
        if lib.check_poses(robot, sc, q).mean() >= min_free:
            return sc
    raise RuntimeError(f"no feasible scene for {robot} from seed {seed}")


def _problem(robot, make, seed):
    rng = np.random.default_rng(seed)
    sc = feasible_scene(robot, make, seed, rng)
    q = free_poses(robot, sc, 2, rng)
    return {"robot": robot, "start": q[0], "goals": q[1][None], "scene": sc}


def c0_problem(seed=0):
    return _problem("panda", lambda s: S.scene_spheres(s, 12, avoid=fixed_spheres("panda")), seed)


def c1_scene(seed=1):
    rng = np.random.default_rng(seed)
    return feasible_scene("panda", lambda s: S.scene_obb_cylinders(s, 24, 8, avoid=fixed_spheres("panda")), seed,
                          rng, min_free=0.02)


def c1_problems(n, seed=1):
    sc = c1_scene(seed)
    rng = np.random.default_rng(seed + 7)
    q = free_poses("panda", sc, 2 * n, rng)
    return [{"robot": "panda", "start": q[2 * i], "goals": q[2 * i + 1][None], "scene": sc} for i in range(n)]


def c1_edges(n=1_000_000, seed=0, max_len=0.5):
    """Edges (a, b): a uniform in the Panda limits, b = a + L u with u uniform
    on the unit sphere and L ~ U(0, max_len], redrawn while b leaves the limits."""
    lo, span = limits("panda")
    lo = lo.astype(np.float64)
    hi = lo + span
    rng = np.random.default_rng(seed)
    d = len(lo)
    a = lo + (hi - lo) * rng.random((n, d))
    b = np.empty_like(a)
    todo = np.arange(n)
    while len(todo):
        u = rng.normal(size=(len(todo), d))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        L = max_len * (1.0 - rng.random((len(todo), 1)))
        cand = a[todo] + L * u
        ok = np.all((cand >= lo) & (cand <= hi), axis=1)
        b[todo[ok]] = cand[ok]
        todo = todo[~ok]
    return np.ascontiguousarray(np.stack([a, b], 1).astype(np.float32))


def c2_problems(n=256, seed=0):
    avoid = fixed_spheres("fetch")
    return [_problem("fetch", lambda s: S.scene_mixed(s, 8, 64, avoid=avoid), seed + i) for i in range(n)]


def c3_problems(n_baxter=128, n_panda=256, n_fetch=128, seed=0):
    out = []
    for robot, n in (("baxter", n_baxter), ("panda", n_panda), ("fetch", n_fetch)):
        avoid = fixed_spheres(robot)
        for i in range(n):
            out.append(_problem(robot, lambda s: S.scene_shelf(s, 4, 128, avoid=avoid), seed + 1000 * len(out) + i))
    return out


def ur5_scene(seed):
    """6 spheres, 4 oriented cuboids and 2 capsules in the UR5's reach, none
    within 5 cm of its base sphere."""
    rng = np.random.default_rng(seed)
    base = fixed_spheres("ur5")

    def clear(c, r):
        return bool(np.all(np.linalg.norm(base[:, :3] - c, axis=1) > base[:, 3] + r + 0.05))
    sph, cub, cap = [], [], []
    while len(sph) < 6:
        c, r = np.r_[rng.uniform(-0.8, 0.8, 2), rng.uniform(0.0, 0.9)], rng.uniform(0.05, 0.12)
        if clear(c, r):
            sph.append(S.sphere(c, r))
    while len(cub) < 4:
        c, h = np.r_[rng.uniform(-0.8, 0.8, 2), rng.uniform(0.0, 0.8)], rng.uniform(0.03, 0.12, 3)
        if clear(c, float(np.linalg.norm(h))):
            cub.append(S.cuboid(c, rng.uniform(-np.pi, np.pi, 3), h))
    while len(cap) < 2:
        c = np.r_[rng.uniform(-0.8, 0.8, 2), rng.uniform(0.1, 0.8)]
        if clear(c, 0.2):
            cap.append(S.cylinder_center(c, rng.uniform(-np.pi, np.pi, 3), 0.05, 0.3))
    return S.Scene.build(spheres=sph, capsules=cap, cuboids=cub)


def ur5_problems(n, seed=0):
    return [_problem("ur5", ur5_scene, seed + 1000 * i) for i in range(n)]


def serving_request(i, seed=0):
    """Robot Panda : Fetch : Baxter = 2 : 1 : 1, a fresh 4-128 primitive scene."""
    rng = np.random.default_rng(seed * 1000003 + i)
    robot = ("panda", "panda", "fetch", "baxter")[int(rng.integers(4))]
    avoid = fixed_spheres(robot)
    return _problem(robot, lambda s: S.scene_serving(s, avoid=avoid), seed * 1000003 + i)
