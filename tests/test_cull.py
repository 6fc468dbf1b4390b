"""The bounding-ball culls of the validation kernels are conservative, and
self-collision groups are dealt to exactly one wave.

The kernels skip a cuboid (capsule) test when the robot sphere's centre is
farther than r + cull from the box centre (segment midpoint), with cull =
half-diagonal (|v|/2 + capsule radius) padded by 2^-16 relative and 1 mm.
Skipping is exact only if every skipped test value is >= 0; these tests
evaluate the reference's squared forms (src/planning/utils.cuh:113-175) in
float32 with the kernels' association on points just beyond the cull
radius, where the ball is tightest (box corners, capsule ends)."""
import os
import re

import numpy as np

from prrtc_amd import scenes as S

F32 = np.float32


def _cuboid_value(cub, p, r):
    """utils.cuh:158-175 in float32, fmaf association of the kernel."""
    c, a, h = cub[0:3], cub[3:12].reshape(3, 3), cub[12:15]
    d = (p - c).astype(F32)
    acc = F32(0)
    es = []
    for i in range(3):
        proj = F32(F32(a[i, 0] * d[..., 0]) + F32(a[i, 1] * d[..., 1])) + F32(a[i, 2] * d[..., 2])
        es.append(np.maximum(F32(0), np.abs(proj) - h[i]).astype(F32))
    acc = (es[0] * es[0] + es[1] * es[1] + es[2] * es[2]).astype(F32)
    return (acc - F32(r) * F32(r)).astype(F32)


def _capsule_value(cap, p, r):
    """utils.cuh:113-126 in float32."""
    p1, v, rc, rdv = cap[0:3], cap[3:6], cap[6], cap[7]
    d = (p - p1).astype(F32)
    dot = (d[..., 0] * v[0] + d[..., 1] * v[1] + d[..., 2] * v[2]).astype(F32)
    cdf = np.clip(dot * rdv, F32(0), F32(1)).astype(F32)
    q = (p1 + cdf[..., None] * v).astype(F32)
    e = (p - q).astype(F32)
    rs = F32(r + rc)
    return ((e * e).sum(-1) - rs * rs).astype(F32)


def test_cuboid_cull_is_conservative():
    rng = np.random.default_rng(5)
    worst = np.inf
    for _ in range(200):
        h = rng.uniform(0.005, 0.6, 3).astype(F32)
        c = rng.uniform(-1, 1, 3).astype(F32)
        cub = S.cuboid(c, rng.uniform(-3, 3, 3), h)
        cull = F32(np.linalg.norm(h.astype(np.float64)) * (1 + 2.0 ** -16) + 1e-3)
        r = F32(rng.uniform(0.01, 0.2))
        R = cub[3:12].reshape(3, 3).T
        # directions: the 8 corner diagonals (tight) and random ones
        signs = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], F32)
        dirs = np.concatenate([(R @ (signs * h).T).T, rng.normal(size=(64, 3))]).astype(np.float64)
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        dist = (r + cull) * (1 + 1e-6) + rng.uniform(0, 1e-4, (len(dirs), 1))
        p = (c + dirs * dist).astype(F32)
        far = np.sum((p - c).astype(F32) ** 2, -1) > F32(r + cull) ** 2
        v = _cuboid_value(cub, p, r)
        assert np.all(v[far] >= 0)
        worst = min(worst, float(v[far].min()) if far.any() else np.inf)
    assert worst > 0


def test_capsule_cull_is_conservative():
    rng = np.random.default_rng(6)
    for _ in range(200):
        p1 = rng.uniform(-1, 1, 3).astype(F32)
        p2 = (p1 + rng.uniform(-0.8, 0.8, 3)).astype(F32)
        rc = F32(rng.uniform(0.01, 0.3))
        cap = S.capsule_endpoints(p1, p2, rc)
        v = cap[3:6].astype(np.float64)
        m = (p1 + 0.5 * v).astype(F32)
        cull = F32((0.5 * np.linalg.norm(v) + rc) * (1 + 2.0 ** -16) + 1e-3)
        r = F32(rng.uniform(0.01, 0.2))
        axis = v / np.linalg.norm(v)
        dirs = np.concatenate([axis[None], -axis[None], rng.normal(size=(64, 3))])
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        p = (m + dirs * ((r + cull) * (1 + 1e-6))).astype(F32)
        far = np.sum((p - m).astype(F32) ** 2, -1) > F32(r + cull) ** 2
        assert np.all(_capsule_value(cap, p, r)[far] >= 0)


def test_self_groups_dealt_once():
    """Generated self groups are numbered 0..n-1 in check order, and the
    kernels' 12-group ownership pattern gives each to exactly one wave for
    every task width 1..4."""
    here = os.path.join(os.path.dirname(__file__), "..", "prrtc_amd", "csrc", "robots")
    for robot in ("panda", "fetch", "baxter"):
        text = open(os.path.join(here, robot + ".gen.h")).read()
        js = [int(j) for j in re.findall(r"ctx\.mine\((\d+)\)", text)]
        assert js == list(range(len(js))) and len(js) > 0
        for W in (1, 2, 3, 4):
            masks = [sum(1 << i for i in range(12) if i % W == w) for w in range(W)]
            owners = [[w for w in range(W) if (masks[w] >> (j % 12)) & 1] for j in js]
            assert all(len(o) == 1 and o[0] == j % W for o, j in zip(owners, js))


def test_group_bounds_hold_every_leaf(oracle_robots):
    """Every bounded leaf group of the generated check sequences: on random
    poses (oracle FK), each leaf sphere lies inside the emitted ball about
    the gate centre with most of the 1 mm pad to spare, so skipping an
    obstacle that stays outside the ball (plus its own bound) cannot change
    a leaf test."""
    from prrtc_amd.robots import limits, load_model
    here = os.path.join(os.path.dirname(__file__), "..", "prrtc_amd", "csrc", "robots")
    rng = np.random.default_rng(9)
    for robot in ("panda", "fetch", "baxter"):
        m = load_model(robot)
        by_label = {c["label"]: c for c in m["env_checks"]}
        text = open(os.path.join(here, robot + ".gen.h")).read()
        blocks = re.findall(r"/\* env leaves: (.*?) \*/(.*?)(?=/\* env leaves: |/\* self: |\Z)", text, re.S)
        lo, span = limits(robot)
        q = (lo + span * rng.random((2000, len(lo)), dtype=np.float32)).astype(np.float32)
        c = np.asarray(oracle_robots[robot].fk(q))
        r = np.array([s["r"] for s in m["spheres"]])
        bounded = set()
        for label, body in blocks:
            mb = re.search(r"env_any_bounded<\d+, true>\(.*, (-?[0-9.e+-]+)f\);", body)
            if not mb:
                continue
            bounded.add(label)
            gr = float(mb.group(1))
            chk = by_label[label]
            g = chk["gate"]
            reach = max(float((np.linalg.norm(c[:, l] - c[:, g], axis=1) + r[l]).max()) for l in chk["leaves"])
            assert reach <= gr - 0.9e-3, (robot, label, reach, gr)
        # exactly the checks of at least 6 leaves that all ride on the gate's frame
        sph = m["spheres"]
        want = {c["label"] for c in m["env_checks"] if not c["direct"] and len(c["leaves"]) >= 6 and
                all(sph[l]["frame"] == sph[c["gate"]]["frame"] for l in c["leaves"])}
        assert bounded == want, (robot, bounded ^ want)
