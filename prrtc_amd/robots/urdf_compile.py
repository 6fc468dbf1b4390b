"""Robot-model compiler: a URDF subset plus a sphere list -> a model table.

    python -m prrtc_amd.robots.urdf_compile robot.urdf spheres.json -o robot.json

The model table (``<robot>.json``, the input of ``gen_headers``) describes FK
as a chain of joint frames, T_j = T_parent(j) * F_j * Rz(q_j) for a revolute
joint or * Tz(q_j) for a prismatic one, with every collision sphere given by
its frame j (-1: the base) and its centre in that frame.  This module builds
that table from the robot's description, the way the reference's robot code
was produced from a URDF and a sphere set (the pattern of
src/planning/utils.cuh:326-33131 and src/planning/Robots.hh):

URDF subset
    <robot>, <link name>, <joint name type> with type revolute, continuous,
    prismatic or fixed, <parent link>, <child link>, <origin xyz rpy>
    (rpy = fixed-axis roll-pitch-yaw: R = Rz(yaw) Ry(pitch) Rx(roll)),
    <axis xyz> (default 1 0 0) and <limit lower upper>.  A continuous joint
    without limits spans [-pi, pi].  Mimic joints, floating and planar joints
    are not part of the subset.

Frames
    The actuated joints, in the order they appear in the file (or the order
    the sphere file's "joints" names), are the configuration's coordinates.
    Joint k's frame has its z axis along the joint axis: with R_a the
    smallest rotation taking z to the axis, URDF's
    T_child = T_parent_link * Origin * Motion(axis, q) equals
    J_k(q) * R_a^T with J_k(q) = T_parent_link * Origin * R_a * Rz(q) (or
    Tz(q)).  So F_k = A * Origin * R_a, where A is the fixed transform from
    the nearest actuated ancestor's frame to the parent link (fixed joints
    folded in), and a sphere on link L sits at A_L * p in its frame.

Sphere file (JSON)
    {"name": "...",
     "joints": [names]                    optional DoF order
     "limits": {"lower": [...], "range": [...]}   optional override
     "spheres": [{"link": L, "xyz": [x, y, z], "r": r}, ...],   link frames
     "env": [{"label": s, "gate": i, "leaves": [..]} | {"label": s, "sphere": i}],
     "self": [{"label": s, "gate": [i, j], "pairs": [[a, b], ..]} | {"label": s, "pair": [i, j]}],
     "order": [["env", k] | ["self", k], ...]}            optional
    Indices refer to "spheres".  A check group tests its gate sphere first
    and its leaves (or sphere pairs) only where the gate hits; "sphere" /
    "pair" entries are single tests.  Without "order" the environment groups
    run first, in the listed order, then the self groups.  A group may give
    "leaves": "auto" and no gate: the gate is then added as a ball about the
    leaves' centroid that holds every leaf (see ``bounding_gate``).

``export_spheres`` goes the other way for an existing table: it expresses
each sphere in its URDF link frame (zero pose), so a fitted table can be
restated as a URDF plus a sphere file and compiled again.
"""
from __future__ import annotations

import argparse
import json
import math
import xml.etree.ElementTree as ET

import numpy as np

ACTUATED = ("revolute", "continuous", "prismatic")


def _floats(text, n, default):
    if text is None:
        return list(default)
    v = [float(t) for t in text.split()]
    if len(v) != n:
        raise ValueError(f"expected {n} numbers, got {text!r}")
    return v


def rpy_matrix(roll, pitch, yaw):
    """URDF's fixed-axis roll-pitch-yaw: Rz(yaw) @ Ry(pitch) @ Rx(roll)."""
    cr, sr = math.cos(roll), math.sin(roll)
    cp, sp = math.cos(pitch), math.sin(pitch)
    cy, sy = math.cos(yaw), math.sin(yaw)
    rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return rz @ ry @ rx


def align_z(axis):
    """Smallest rotation R with R @ z = axis (unit)."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    z = np.array([0.0, 0.0, 1.0])
    v = np.cross(z, a)
    c = float(a[2])
    if np.linalg.norm(v) < 1e-12:
        return np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])   # 180 degrees about x
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + vx + vx @ vx / (1.0 + c)


def _h(R=None, t=None):
    T = np.eye(4)
    if R is not None:
        T[:3, :3] = R
    if t is not None:
        T[:3, 3] = t
    return T


def parse_urdf(text: str) -> dict:
    """Links, joints and the root link of a URDF-subset document."""
    root = ET.fromstring(text)
    if root.tag != "robot":
        raise ValueError("not a URDF <robot> document")
    links = [l.get("name") for l in root.findall("link")]
    joints = []
    for j in root.findall("joint"):
        jt = j.get("type")
        if jt not in ACTUATED + ("fixed",):
            raise ValueError(f"joint {j.get('name')}: type {jt!r} is outside the supported subset")
        if j.find("mimic") is not None:
            raise ValueError(f"joint {j.get('name')}: mimic joints are outside the supported subset")
        o = j.find("origin")
        ax = j.find("axis")
        lim = j.find("limit")
        joints.append({
            "name": j.get("name"), "type": jt,
            "parent": j.find("parent").get("link"), "child": j.find("child").get("link"),
            "xyz": _floats(o.get("xyz") if o is not None else None, 3, (0, 0, 0)),
            "rpy": _floats(o.get("rpy") if o is not None else None, 3, (0, 0, 0)),
            "axis": _floats(ax.get("xyz") if ax is not None else None, 3, (1, 0, 0)),
            "lower": float(lim.get("lower")) if lim is not None and lim.get("lower") is not None else None,
            "upper": float(lim.get("upper")) if lim is not None and lim.get("upper") is not None else None,
        })
    children = {j["child"] for j in joints}
    roots = [l for l in links if l not in children]
    if len(roots) != 1:
        raise ValueError(f"expected one root link, found {roots}")
    if len(children) != len(joints):
        raise ValueError("a link is the child of two joints")
    return {"name": root.get("name"), "links": links, "joints": joints, "root": roots[0]}


def link_frames(urdf: dict, order=None):
    """Walk the tree from the root.  Returns (actuated joints in DoF order,
    {link: (frame index or -1, 4x4 transform from that frame to the link)},
    [F_k 4x4], [parent frame of k])."""
    by_parent = {}
    for j in urdf["joints"]:
        by_parent.setdefault(j["parent"], []).append(j)
    act = [j for j in urdf["joints"] if j["type"] in ACTUATED]
    names = [j["name"] for j in act]
    if order is not None:
        missing = set(order) ^ set(names)
        if missing:
            raise ValueError(f"joint order does not match the actuated joints: {sorted(missing)}")
        names = list(order)
    index = {n: k for k, n in enumerate(names)}
    F = [None] * len(names)
    parent = [None] * len(names)
    frames = {urdf["root"]: (-1, np.eye(4))}
    stack = [urdf["root"]]
    while stack:
        link = stack.pop()
        a, A = frames[link]
        for j in by_parent.get(link, []):
            origin = _h(rpy_matrix(*j["rpy"]), j["xyz"])
            if j["type"] == "fixed":
                frames[j["child"]] = (a, A @ origin)
            else:
                k = index[j["name"]]
                Ra = align_z(j["axis"])
                F[k] = A @ origin @ _h(Ra)
                parent[k] = a
                frames[j["child"]] = (k, _h(Ra.T))
            stack.append(j["child"])
    if any(p is not None and p >= k for k, p in enumerate(parent)):
        raise ValueError("the DoF order lists a joint before its actuated ancestor")
    return [next(j for j in act if j["name"] == n) for n in names], frames, F, parent


def compile_model(urdf_text: str, spec: dict) -> dict:
    """Model table (the schema of prrtc_amd/robots/<robot>.json)."""
    urdf = parse_urdf(urdf_text)
    act, frames, F, parent = link_frames(urdf, spec.get("joints"))
    lower, rng = [], []
    for j in act:
        lo, hi = j["lower"], j["upper"]
        if lo is None or hi is None:
            if j["type"] != "continuous":
                raise ValueError(f"joint {j['name']} has no limits")
            lo, hi = -math.pi, math.pi
        lower.append(lo)
        rng.append(hi - lo)
    if "limits" in spec:
        lower, rng = list(spec["limits"]["lower"]), list(spec["limits"]["range"])
        if len(lower) != len(act) or len(rng) != len(act):
            raise ValueError("limits override must give one value per actuated joint")
    joints = [{"parent": int(parent[k]), "type": "prismatic" if j["type"] == "prismatic" else "revolute",
               "rot": F[k][:3, :3].tolist(), "trans": F[k][:3, 3].tolist()} for k, j in enumerate(act)]
    spheres = []
    for s in spec["spheres"]:
        if s["link"] not in frames:
            raise ValueError(f"sphere on unknown link {s['link']!r}")
        a, A = frames[s["link"]]
        p = A @ np.array(list(s["xyz"]) + [1.0])
        spheres.append({"frame": int(a), "local": p[:3].tolist(), "r": float(s["r"]),
                        "label": s.get("label", s["link"])})
    env, selfc = [], []
    for g in spec.get("env", []):
        if "sphere" in g:
            env.append({"label": g["label"], "gate": int(g["sphere"]), "leaves": [], "direct": True})
        else:
            gate = g.get("gate")
            leaves = [int(i) for i in g["leaves"]]
            if gate is None:
                gate = len(spheres)
                spheres.append(bounding_gate(spheres, leaves, g["label"]))
            env.append({"label": g["label"], "gate": int(gate), "leaves": leaves, "direct": False})
    for g in spec.get("self", []):
        if "pair" in g:
            selfc.append({"label": g["label"], "gate": [int(v) for v in g["pair"]], "pairs": [], "direct": True})
        else:
            selfc.append({"label": g["label"], "gate": [int(v) for v in g["gate"]],
                          "pairs": [[int(a), int(b)] for a, b in g["pairs"]], "direct": False})
    order = spec.get("order") or [["env", k] for k in range(len(env))] + [["self", k] for k in range(len(selfc))]
    for kind, k in order:
        if not 0 <= k < (len(env) if kind == "env" else len(selfc)):
            raise ValueError(f"order names {kind} {k}, which does not exist")
    return {"name": spec.get("name", urdf["name"]), "dof": len(act), "lower": lower, "range": rng,
            "joints": joints, "spheres": spheres, "env_checks": env, "self_checks": selfc,
            "order": [list(o) for o in order]}


def bounding_gate(spheres, leaves, label):
    """A ball about the leaves' centroid that holds every leaf, on the
    leaves' common frame (they must share one)."""
    frames = {spheres[i]["frame"] for i in leaves}
    if len(frames) != 1:
        raise ValueError(f"{label}: an automatic gate needs leaves on one frame")
    P = np.array([spheres[i]["local"] for i in leaves])
    c = P.mean(0)
    r = max(float(np.linalg.norm(P[i] - c)) + spheres[k]["r"] for i, k in enumerate(leaves))
    return {"frame": frames.pop(), "local": c.tolist(), "r": r, "label": label}


def model_fk(model: dict, q) -> np.ndarray:
    """Sphere centres [n, S, 3] of a model table in binary64 (the FK the
    generated device code evaluates in binary32)."""
    q = np.atleast_2d(np.asarray(q, dtype=np.float64))
    n = len(q)
    R = {-1: np.broadcast_to(np.eye(3), (n, 3, 3))}
    t = {-1: np.zeros((n, 3))}
    for k, jt in enumerate(model["joints"]):
        p = jt["parent"]
        Fr, Ft = np.array(jt["rot"]), np.array(jt["trans"])
        M = R[p] @ Fr
        tt = t[p] + R[p] @ Ft
        if jt["type"] == "revolute":
            c, s = np.cos(q[:, k]), np.sin(q[:, k])
            Rz = np.zeros((n, 3, 3))
            Rz[:, 0, 0], Rz[:, 0, 1], Rz[:, 1, 0], Rz[:, 1, 1], Rz[:, 2, 2] = c, -s, s, c, 1.0
            R[k], t[k] = M @ Rz, tt
        else:
            R[k], t[k] = M, tt + M[:, :, 2] * q[:, k:k + 1]
    out = np.zeros((n, len(model["spheres"]), 3))
    for i, s in enumerate(model["spheres"]):
        f = s["frame"]
        out[:, i] = t[f] + R[f] @ np.array(s["local"])
    return out


def urdf_link_poses(urdf_text: str, q, order=None) -> dict:
    """World transform of every link at configuration q (one pose), by
    URDF's own composition rule, for checks and ``export_spheres``."""
    urdf = parse_urdf(urdf_text)
    act, _, _, _ = link_frames(urdf, order)
    qv = {j["name"]: float(v) for j, v in zip(act, np.asarray(q, dtype=np.float64))}
    by_parent = {}
    for j in urdf["joints"]:
        by_parent.setdefault(j["parent"], []).append(j)
    out = {urdf["root"]: np.eye(4)}
    stack = [urdf["root"]]
    while stack:
        link = stack.pop()
        for j in by_parent.get(link, []):
            T = out[link] @ _h(rpy_matrix(*j["rpy"]), j["xyz"])
            if j["type"] in ("revolute", "continuous"):
                ax = np.asarray(j["axis"]) / np.linalg.norm(j["axis"])
                K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
                th = qv[j["name"]]
                T = T @ _h(np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K)   # Rodrigues
            elif j["type"] == "prismatic":
                ax = np.asarray(j["axis"]) / np.linalg.norm(j["axis"])
                T = T @ _h(t=ax * qv[j["name"]])
            out[j["child"]] = T
            stack.append(j["child"])
    return out


def export_spheres(model: dict, urdf_text: str, link_of_label=None) -> dict:
    """A sphere file for `model` against `urdf_text`: every sphere expressed
    in the frame of the URDF link its label names (or link_of_label(label)),
    from both FKs at the zero pose; the check structure and order copied."""
    zero = np.zeros(model["dof"])
    world = model_fk(model, zero)[0]
    poses = urdf_link_poses(urdf_text, zero)
    spheres = []
    for i, s in enumerate(model["spheres"]):
        link = link_of_label(s["label"]) if link_of_label else s["label"]
        if link not in poses:
            raise ValueError(f"sphere {i}: label {s['label']!r} names no URDF link")
        T = poses[link]
        local = T[:3, :3].T @ (world[i] - T[:3, 3])
        spheres.append({"link": link, "xyz": local.tolist(), "r": s["r"], "label": s["label"]})
    env = [{"label": c["label"], "sphere": c["gate"]} if c["direct"] else
           {"label": c["label"], "gate": c["gate"], "leaves": c["leaves"]} for c in model["env_checks"]]
    selfc = [{"label": c["label"], "pair": c["gate"]} if c["direct"] else
             {"label": c["label"], "gate": c["gate"], "pairs": c["pairs"]} for c in model["self_checks"]]
    return {"name": model["name"], "limits": {"lower": model["lower"], "range": model["range"]},
            "spheres": spheres, "env": env, "self": selfc, "order": model["order"]}


def main(argv=None):
    ap = argparse.ArgumentParser(description="compile a URDF subset and a sphere file into a model table")
    ap.add_argument("urdf")
    ap.add_argument("spheres")
    ap.add_argument("-o", "--out", required=True)
    args = ap.parse_args(argv)
    with open(args.urdf) as f:
        urdf_text = f.read()
    with open(args.spheres) as f:
        spec = json.load(f)
    model = compile_model(urdf_text, spec)
    with open(args.out, "w") as f:
        json.dump(model, f, indent=1)
    print(f"{args.out}: {model['dof']} joints, {len(model['spheres'])} spheres, "
          f"{len(model['env_checks'])} env groups, {len(model['self_checks'])} self groups")


if __name__ == "__main__":
    main()
