"""Algorithmic work of collision checking, counted the way the reference
executes it -- the denominator of the validation roofline.

For one pose the reference's fkcc<Robot> (src/planning/utils.cuh:336-33131)
runs its straight-line FK and then, in check order, every link's gate
sphere against the obstacle tables (each table loop stopping at its first
hit, utils.cuh:211-242), every leaf under a gate that hit, every
self-collision gate pair and, under a pair that hit, its sphere pairs.  For a
collision-free pose nothing stops early except those per-gate obstacle loops,
so its work is

    F_FK + sum over executed tests of w(test type)

with F_FK the reference's FK multiply/add/subtract count (half-angle scalings
excluded) and the test weights below (add, sub, mul, min, max, abs, compare
= 1 flop, FMA = 2).  The counts are taken from sphere centres produced by the
HIP FK kernel (prrtc_fk_spheres) on a sample of the poses.
"""
from __future__ import annotations

import numpy as np

from .robots import load_model

# reference FK op counts (mul + add + sub of utils.cuh's generated code)
F_FK = {"sphere": 0, "panda": 1079, "fetch": 962, "baxter": 1750}
# flops per sphere-vs-obstacle test
W = {"sphere": 12, "capsule": 29, "z_capsule": 19, "cuboid": 34, "z_cuboid": 25, "self": 12}
TABLES = ("spheres", "capsules", "z_capsules", "cuboids", "z_cuboids")
TABLE_W = (W["sphere"], W["capsule"], W["z_capsule"], W["cuboid"], W["z_cuboid"])


def _hits(scene, c, r):
    """Per table: hit matrix [n_poses, n_obstacles] of spheres (c [n,3], r)."""
    out = []
    x, y, z = c[:, 0:1].astype(np.float64), c[:, 1:2].astype(np.float64), c[:, 2:3].astype(np.float64)
    for name in TABLES:
        t = np.asarray(getattr(scene, name), dtype=np.float64)
        if len(t) == 0:
            out.append(np.zeros((len(c), 0), dtype=bool))
            continue
        if name == "spheres":
            d2 = (t[:, 0] - x) ** 2 + (t[:, 1] - y) ** 2 + (t[:, 2] - z) ** 2
            out.append(d2 < (t[:, 3] + r) ** 2)
        elif name in ("capsules", "z_capsules"):
            px, py, pz = x - t[:, 0], y - t[:, 1], z - t[:, 2]
            cdf = np.clip((px * t[:, 3] + py * t[:, 4] + pz * t[:, 5]) * t[:, 7], 0, 1)
            d2 = (px - t[:, 3] * cdf) ** 2 + (py - t[:, 4] * cdf) ** 2 + (pz - t[:, 5] * cdf) ** 2
            out.append(d2 < (t[:, 6] + r) ** 2)
        else:
            xs, ys, zs = x - t[:, 0], y - t[:, 1], z - t[:, 2]
            acc = 0
            for a in range(3):
                proj = xs * t[:, 3 + 3 * a] + ys * t[:, 4 + 3 * a] + zs * t[:, 5 + 3 * a]
                acc = acc + np.maximum(0, np.abs(proj) - t[:, 12 + a]) ** 2
            out.append(acc < r * r)
    return out


def _env_flops(scene, c, r, early_exit):
    """Flops of sphere_environment_in_collision on each pose's sphere; with
    early_exit the loops stop at the first hit, table order."""
    hits = _hits(scene, c, r)
    n = len(c)
    flops = np.zeros(n)
    done = np.zeros(n, dtype=bool)
    for h, w in zip(hits, TABLE_W):
        m = h.shape[1]
        if m == 0:
            continue
        if early_exit:
            first = np.where(h.any(1), h.argmax(1) + 1, m)
            flops += np.where(done, 0, first * w)
            done |= h.any(1)
        else:
            flops += m * w
    return flops, np.logical_or.reduce([h.any(1) for h in hits if h.shape[1]] or [np.zeros(n, bool)])


def pose_flops(robot: str, scene, centres: np.ndarray) -> np.ndarray:
    """Reference flops per collision-free pose; centres [n, S, 3] from the
    robot's FK (sphere order of the model table)."""
    n = centres.shape[0]
    if robot == "sphere":
        f, _ = _env_flops(scene, centres[:, 0], 1.0, True)
        return f
    m = load_model(robot)
    rad = [s["r"] for s in m["spheres"]]
    flops = np.full(n, float(F_FK[robot]))
    for chk in m["env_checks"]:
        g = chk["gate"]
        f, hit = _env_flops(scene, centres[:, g], rad[g], True)
        flops += f
        if chk["direct"]:
            continue
        for l in chk["leaves"]:
            fl, _ = _env_flops(scene, centres[:, l], rad[l], False)
            flops += np.where(hit, fl, 0.0)
    for chk in m["self_checks"]:
        a, b = chk["gate"]
        flops += W["self"]
        d2 = ((centres[:, a].astype(np.float64) - centres[:, b]) ** 2).sum(1)
        hit = d2 < (rad[a] + rad[b]) ** 2
        if not chk["direct"]:
            flops += np.where(hit, W["self"] * len(chk["pairs"]), 0.0)
    return flops


def gate_pass_flops(robot: str, scene) -> float:
    """Algorithmic work per pose of a planner validation pass, as an
    estimate: the reference's FK plus every env gate sphere tested against
    every obstacle of the scene (utils.cuh:211-242 for each gate).  Leaves
    under hit gates and self-collision pairs are not counted, and the gate
    loops' early exits at a hit are not taken off; the planner's colliding
    passes stop before most of this work, its free passes do more."""
    per_obstacle = sum(len(getattr(scene, t)) * w for t, w in zip(TABLES, TABLE_W))
    if robot == "sphere":
        return float(per_obstacle)
    m = load_model(robot)
    return float(F_FK[robot] + len(m["env_checks"]) * per_obstacle)
