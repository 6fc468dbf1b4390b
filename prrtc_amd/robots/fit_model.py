"""Fit table-driven robot models from the reference's generated FK code.

Offline tool, run once in a container that holds the reference sources:

    python -m prrtc_amd.robots.fit_model /path/to/reference

For each robot it parses ``fkcc<Robot>`` (src/planning/utils.cuh) with
``fkcc_text``, evaluates every sphere centre in binary64 at probe poses and
recovers a kinematic tree:

* each joint's screw axis at the zero pose, from a rigid (Kabsch) fit of the
  spheres that move when only that joint moves (revolute: rotation axis and a
  point on it; prismatic: translation direction);
* each joint's parent, from which joints the moving spheres depend on;
* per joint a fixed transform ``F_j`` (parent joint frame -> this joint's
  frame, z along the axis), so that FK is
  ``T_j = T_parent(j) * F_j * Joint_j(q_j)`` (Joint = Rz or Tz);
* every sphere's frame and local offset, and the collision-check structure
  (link gate spheres, leaf spheres, self-collision gate pairs and sphere pairs)
  in the reference's check order.

Joint limits come from src/planning/Robots.hh (``get_s_m`` = range,
``get_s_a`` = lower bound).  The output JSON holds numbers and structure only.
The fitted model is checked here against the parsed program at random poses
(binary64, worst error printed) before it is written.
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

from . import fkcc_text as F

HERE = os.path.dirname(os.path.abspath(__file__))
ROBOT_STRUCT = {"panda": "Panda", "fetch": "Fetch", "baxter": "Baxter"}
DOF = {"panda": 7, "fetch": 8, "baxter": 14}


def read_limits(robots_hh: str, robot: str):
    text = open(robots_hh).read()
    start = text.index("struct " + ROBOT_STRUCT[robot])
    sub = text[start:]
    out = {}
    for fn in ("get_s_m", "get_s_a"):
        i = sub.index(fn)
        body = sub[sub.index("{", sub.index("values[]", i)) + 1: sub.index("}", sub.index("values[]", i))]
        out[fn] = [float(v.strip().rstrip("f")) for v in body.split(",") if v.strip()]
    return out["get_s_a"], out["get_s_m"]


def _kabsch(P, Q):
    """Rigid transform (R, t) minimising |R P + t - Q| (points as rows)."""
    pc, qc = P.mean(0), Q.mean(0)
    H = (P - pc).T @ (Q - qc)
    U, _S, Vt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(Vt.T @ U.T))
    D = np.diag([1.0, 1.0, d])
    R = Vt.T @ D @ U.T
    return R, qc - R @ pc


def _min_rotation(a, b):
    """Smallest rotation taking unit vector a to unit vector b."""
    v = np.cross(a, b)
    c = float(np.dot(a, b))
    if np.linalg.norm(v) < 1e-14:
        if c > 0:
            return np.eye(3)
        # 180 degrees about any axis orthogonal to a
        ortho = np.array([1.0, 0, 0]) if abs(a[0]) < 0.9 else np.array([0, 1.0, 0])
        ax = np.cross(a, ortho)
        ax /= np.linalg.norm(ax)
        return 2 * np.outer(ax, ax) - np.eye(3)
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + vx + vx @ vx * (1.0 / (1.0 + c))


def _snap(x, tol=1e-9):
    x = np.array(x, dtype=np.float64)
    for val in (0.0, 1.0, -1.0):
        x[np.abs(x - val) < tol] = val
    return x


def _snap_decimal(x, places=5, tol=2e-7):
    """Lengths in the robot description are given to 0.01 mm or coarser; the
    fit recovers them to ~1e-7 m.  Round a value to `places` decimals when it
    lies within `tol` of such a value, keep it otherwise."""
    x = np.array(x, dtype=np.float64)
    r = np.round(x, places)
    return np.where(np.abs(x - r) < tol, r, x)


def fit_robot(ref_root: str, robot: str) -> dict:
    prog = F.parse(os.path.join(ref_root, "src/planning/utils.cuh"), robot)
    deps = F.dependencies(prog)
    dof = DOF[robot]

    spheres, env_checks, self_checks, order = F.structure(prog)

    def sphere_deps(s):
        out = frozenset()
        for a in s["args"][:3]:
            out |= prog._walk(a)
        return out

    for s in spheres:
        s["deps"] = sphere_deps(s)
        r = s["args"][3]
        assert r.kind == "num", "radius must be a literal"
        s["r"] = float(r.value)

    def centres(q):
        return F.sphere_centres(prog, spheres, q, mode="f64")

    home = centres(np.zeros((1, dof)))[0]

    # ---- joint tree -----------------------------------------------------
    moving = {j: [i for i, s in enumerate(spheres) if j in s["deps"]] for j in range(dof)}
    ancestors = {}
    for j in range(dof):
        sets = [spheres[i]["deps"] - {j} for i in moving[j]]
        assert sets, f"{robot}: joint {j} moves no sphere"
        anc = frozenset.intersection(*sets)
        ancestors[j] = anc
    parent = {}
    for j in range(dof):
        anc = ancestors[j]
        parent[j] = max(anc, key=lambda a: len(ancestors[a])) if anc else -1
        if anc:
            assert ancestors[parent[j]] | {parent[j]} == anc, f"{robot}: joint {j} ancestors not a chain"

    joints = []
    G = {-1: (np.eye(3), np.zeros(3))}
    for j in range(dof):
        idx = moving[j]
        P = home[idx]
        thetas = [0.4, -0.7, 1.1]
        Rs, ts = [], []
        for th in thetas:
            q = np.zeros((1, dof))
            q[0, j] = th
            Q = centres(q)[0][idx]
            R, t = _kabsch(P, Q)
            res = np.abs((P @ R.T + t) - Q).max()
            assert res < 1e-6, f"{robot}: joint {j} motion not rigid (res {res})"
            Rs.append(R)
            ts.append(t)
        rot_angle = [np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1)) for R in Rs]
        if max(rot_angle) < 1e-9:
            jtype = "prismatic"
            axis = ts[0] / thetas[0]
            scale = np.linalg.norm(axis)
            assert abs(scale - 1.0) < 1e-6, f"{robot}: prismatic joint {j} scale {scale}"
            axis = axis / scale
            for th, t in zip(thetas, ts):
                assert np.abs(t - axis * th).max() < 1e-6
            point = G[parent[j]][1].copy()
        else:
            jtype = "revolute"
            R, t, th = Rs[0], ts[0], thetas[0]
            w = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]]) / (2 * np.sin(th))
            w /= np.linalg.norm(w)
            # check every probe angle is a rotation by exactly theta about w
            for th2, R2, t2 in zip(thetas, Rs, ts):
                K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
                Rw = np.eye(3) + np.sin(th2) * K + (1 - np.cos(th2)) * K @ K
                assert np.abs(Rw - R2).max() < 1e-6, f"{robot}: joint {j} angle/axis mismatch"
            # point on the axis: (I - R) p = t, least squares, then the point of
            # the axis line closest to the parent frame origin
            A = np.vstack([np.eye(3) - R2 for R2 in Rs])
            b = np.concatenate(ts)
            p, *_ = np.linalg.lstsq(A, b, rcond=None)
            o = G[parent[j]][1]
            p = p + w * np.dot(o - p, w)
            axis, point = w, p
        Ap, op = G[parent[j]]
        A = _min_rotation(Ap[:, 2], axis) @ Ap
        # F_j = G_parent^-1 G_j.  The generated code carries its rotation
        # constants to ~7 digits, so entries within 1e-6 of 0 / +-1 are the
        # exact values the robot description had; snapping them keeps F_j
        # sparse (cheap FK) and exactly orthonormal.
        Fr = _snap(Ap.T @ A, 1e-6)
        U, _s, Vt = np.linalg.svd(Fr)
        Fr = _snap(U @ Vt, 1e-12)
        Ft = _snap_decimal(_snap(Ap.T @ (point - op), 1e-6))
        # re-derive this joint's world frame from the snapped F_j
        G[j] = (Ap @ Fr, op + Ap @ Ft)
        joints.append({
            "parent": int(parent[j]),
            "type": jtype,
            "rot": Fr.tolist(),
            "trans": Ft.tolist(),
        })

    # ---- sphere frames and local offsets ----------------------------------
    out_spheres = []
    for i, s in enumerate(spheres):
        d = s["deps"]
        if not d:
            frame = -1
        else:
            cands = [j for j in d if ancestors[j] | {j} == d]
            assert len(cands) == 1, f"{robot}: sphere {i} deps {sorted(d)} not a joint path"
            frame = cands[0]
        A, o = G[frame]
        local = A.T @ (home[i] - o)
        out_spheres.append({"frame": int(frame), "local": _snap_decimal(_snap(local, 1e-12)).tolist(), "r": s["r"],
                            "label": s["label"] or ""})

    lower, rng = read_limits(os.path.join(ref_root, "src/planning/Robots.hh"), robot)
    model = {
        "name": robot,
        "dof": dof,
        "lower": lower,
        "range": rng,
        "joints": joints,
        "spheres": out_spheres,
        "env_checks": env_checks,
        "self_checks": self_checks,
        "order": order,
    }
    err = check_model(model, prog, centres)
    print(f"{robot}: {len(out_spheres)} spheres, {len(env_checks)} env checks, "
          f"{len(self_checks)} self checks; worst fit error vs parsed FK (binary64) = {err:.3e} m")
    assert err < 1e-6
    return model


def fk_model(model: dict, q: np.ndarray) -> np.ndarray:
    """Sphere centres of a model table at poses q (binary64), shape (n, S, 3)."""
    q = np.asarray(q, dtype=np.float64)
    n = q.shape[0]
    frames = {-1: (np.broadcast_to(np.eye(3), (n, 3, 3)), np.zeros((n, 3)))}
    for j, jt in enumerate(model["joints"]):
        Rp, tp = frames[jt["parent"]]
        Fr = np.array(jt["rot"])
        Ft = np.array(jt["trans"])
        R = Rp @ Fr
        t = tp + Rp @ Ft
        if jt["type"] == "revolute":
            c, s = np.cos(q[:, j]), np.sin(q[:, j])
            Rz = np.zeros((n, 3, 3))
            Rz[:, 0, 0] = c
            Rz[:, 0, 1] = -s
            Rz[:, 1, 0] = s
            Rz[:, 1, 1] = c
            Rz[:, 2, 2] = 1
            R = R @ Rz
        else:
            t = t + R[:, :, 2] * q[:, j:j + 1]
        frames[j] = (R, t)
    out = np.zeros((n, len(model["spheres"]), 3))
    for i, s in enumerate(model["spheres"]):
        R, t = frames[s["frame"]]
        out[:, i] = t + R @ np.array(s["local"])
    return out


def check_model(model, prog, centres_fn) -> float:
    rng = np.random.default_rng(1234)
    lo = np.array(model["lower"])
    span = np.array(model["range"])
    q = lo + span * rng.random((256, model["dof"]))
    ref = centres_fn(q)
    mine = fk_model(model, q)
    return float(np.abs(ref - mine).max())


def main(argv):
    ref_root = argv[1] if len(argv) > 1 else "/root/reference"
    for robot in ("panda", "fetch", "baxter"):
        model = fit_robot(ref_root, robot)
        with open(os.path.join(HERE, robot + ".json"), "w") as fh:
            json.dump(model, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv)
