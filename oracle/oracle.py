"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module, and only as the checker.  The product
(``prrtc_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAXD = 16
MAXJ = 16

_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)


class Model(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("dof", C.c_int), ("point_radius", C.c_float),
        ("lower", C.c_float * MAXD), ("span", C.c_float * MAXD),
        ("n_joints", C.c_int), ("parent", C.c_int * MAXJ), ("prismatic", C.c_int * MAXJ),
        ("rot", (C.c_float * 9) * MAXJ), ("trans", (C.c_float * 3) * MAXJ),
        ("n_spheres", C.c_int), ("frame", _i32p), ("local", _f32p), ("radius", _f32p),
        ("n_env", C.c_int), ("env_gate", _i32p), ("env_direct", _i32p), ("env_leaf_off", _i32p), ("env_leaf", _i32p),
        ("n_self", C.c_int), ("self_ga", _i32p), ("self_gb", _i32p), ("self_direct", _i32p),
        ("self_pair_off", _i32p), ("self_pa", _i32p), ("self_pb", _i32p),
        ("n_order", C.c_int), ("order_kind", _i32p), ("order_idx", _i32p),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("spheres", _f32p), ("n_spheres", C.c_int),
        ("capsules", _f32p), ("n_capsules", C.c_int),
        ("z_capsules", _f32p), ("n_z_capsules", C.c_int),
        ("cuboids", _f32p), ("n_cuboids", C.c_int),
        ("z_cuboids", _f32p), ("n_z_cuboids", C.c_int),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("max_samples", C.c_int), ("max_iters", C.c_int), ("num_new_configs", C.c_int), ("granularity", C.c_int),
        ("range", C.c_float), ("balance", C.c_int), ("tree_ratio", C.c_float), ("dynamic_domain", C.c_int),
        ("dd_alpha", C.c_float), ("dd_radius", C.c_float), ("dd_min_radius", C.c_float), ("seed", C.c_uint64),
    ]


class Result(C.Structure):
    _fields_ = [
        ("solved", C.c_int), ("start_tree_size", C.c_int), ("goal_tree_size", C.c_int), ("path_length", C.c_int),
        ("iters", C.c_int), ("cost", C.c_float), ("path_nodes", C.c_int),
        ("pose_checks", C.c_int64), ("edge_checks", C.c_int64),
    ]


LIB_V4_PATH = os.path.join(HERE, "liboracle_v4.so")


def build(force: bool = False) -> None:
    src = os.path.join(HERE, "oracle.c")
    stale = [p for p in (LIB_PATH, LIB_V4_PATH)
             if force or not os.path.exists(p) or os.path.getmtime(p) < os.path.getmtime(src)]
    if stale:
        subprocess.check_call(["make", "-s", "-B", "-C", HERE] + [os.path.basename(p) for p in stale])


def host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in line + " "
    except OSError:
        pass
    return False


def use_fastest_build() -> str:
    """Load the AVX-512 build when this host has AVX-512 (the CPU baseline;
    same source and rounding), else the portable one.  Call before any other
    oracle function; returns the build's name."""
    global LIB_PATH
    if _lib is None and host_has_avx512() and os.path.exists(LIB_V4_PATH):
        LIB_PATH = LIB_V4_PATH
    return "x86-64-v4" if LIB_PATH == LIB_V4_PATH else "x86-64-v3"


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.or_fk.argtypes = [C.POINTER(Model), vp, C.c_int, vp]
        L.or_check_poses.argtypes = [C.POINTER(Model), C.POINTER(Scene), vp, C.c_int, vp, vp, vp]
        L.or_validate_edges.argtypes = [C.POINTER(Model), C.POINTER(Scene), vp, C.c_int, C.c_int, vp]
        L.or_validate_edges_mt.argtypes = [C.POINTER(Model), C.POINTER(Scene), vp, C.c_int, C.c_int, vp, C.c_int]
        L.or_plan.argtypes = [C.POINTER(Model), C.POINTER(Scene), vp, vp, C.c_int, C.POINTER(Settings), C.c_int,
                              vp, C.c_int, C.POINTER(Result)]
        L.or_plan.restype = C.c_int
        L.or_rrt_connect.argtypes = [C.POINTER(Model), C.POINTER(Scene), vp, vp, C.c_int, C.c_float, C.c_int,
                                     C.c_uint64, vp, C.c_int, C.POINTER(Result)]
        L.or_rrt_connect.restype = C.c_int
        L.or_revalidate_path.argtypes = [C.POINTER(Model), C.POINTER(Scene), vp, C.c_int, C.c_int]
        L.or_revalidate_path.restype = C.c_int
        L.or_halton_init.argtypes = [vp, C.c_int, C.c_uint64, C.c_int]
        L.or_halton_next.argtypes = [vp, C.c_int, vp]
        L.or_set_fused.argtypes = [C.c_int]
        L.or_set_plant.argtypes = [C.c_int]
        L.or_set_plant.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleRobot:
    """A robot model table loaded into the oracle."""

    def __init__(self, model: dict):
        self.model = model
        self.keep = []
        m = Model()
        m.dof = model["dof"]
        for i, v in enumerate(model["lower"]):
            m.lower[i] = v
        for i, v in enumerate(model["range"]):
            m.span[i] = v
        if model["name"] == "sphere":
            m.kind = 0
            m.point_radius = model["radius"]
            self.n_spheres = 1
        else:
            m.kind = 1
            m.n_joints = len(model["joints"])
            for j, jt in enumerate(model["joints"]):
                m.parent[j] = jt["parent"]
                m.prismatic[j] = 1 if jt["type"] == "prismatic" else 0
                for k, v in enumerate(np.asarray(jt["rot"]).reshape(-1)):
                    m.rot[j][k] = v
                for k, v in enumerate(jt["trans"]):
                    m.trans[j][k] = v
            sp = model["spheres"]
            self.n_spheres = len(sp)
            m.n_spheres = len(sp)
            m.frame = self._i([s["frame"] for s in sp])
            m.local = self._f([s["local"] for s in sp])
            m.radius = self._f([s["r"] for s in sp])
            env = model["env_checks"]
            m.n_env = len(env)
            m.env_gate = self._i([c["gate"] for c in env])
            m.env_direct = self._i([int(c["direct"]) for c in env])
            m.env_leaf_off = self._i(np.concatenate([[0], np.cumsum([len(c["leaves"]) for c in env])]))
            m.env_leaf = self._i([l for c in env for l in c["leaves"]] or [0])
            sf = model["self_checks"]
            m.n_self = len(sf)
            m.self_ga = self._i([c["gate"][0] for c in sf] or [0])
            m.self_gb = self._i([c["gate"][1] for c in sf] or [0])
            m.self_direct = self._i([int(c["direct"]) for c in sf] or [0])
            m.self_pair_off = self._i(np.concatenate([[0], np.cumsum([len(c["pairs"]) for c in sf])]))
            m.self_pa = self._i([p[0] for c in sf for p in c["pairs"]] or [0])
            m.self_pb = self._i([p[1] for c in sf for p in c["pairs"]] or [0])
            order = model["order"]
            m.n_order = len(order)
            m.order_kind = self._i([0 if k == "env" else 1 for k, _ in order])
            m.order_idx = self._i([i for _, i in order])
        self.m = m
        self.dof = m.dof

    def _i(self, v):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(-1))
        self.keep.append(a)
        return a.ctypes.data_as(_i32p)

    def _f(self, v):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1))
        self.keep.append(a)
        return a.ctypes.data_as(_f32p)

    # ---- calls ------------------------------------------------------------
    def fk(self, q):
        q = np.ascontiguousarray(q, dtype=np.float32)
        out = np.zeros((len(q), self.n_spheres, 3), dtype=np.float32)
        lib().or_fk(C.byref(self.m), _ptr(q), len(q), _ptr(out))
        return out

    def check(self, scene, q, margin=False, counts=False):
        q = np.ascontiguousarray(q, dtype=np.float32)
        sc, keep = make_scene(scene)
        free = np.zeros(len(q), dtype=np.uint8)
        mg = np.zeros(len(q), dtype=np.float32) if margin else None
        ct = np.zeros((len(q), 6), dtype=np.int32) if counts else None
        lib().or_check_poses(C.byref(self.m), C.byref(sc), _ptr(q), len(q), _ptr(free),
                             _ptr(mg) if mg is not None else None, _ptr(ct) if ct is not None else None)
        out = [free.astype(bool)]
        if margin:
            out.append(mg)
        if counts:
            out.append(ct)
        return out[0] if len(out) == 1 else tuple(out)

    def validate_edges(self, scene, edges, g, threads=1):
        e = np.ascontiguousarray(edges, dtype=np.float32)
        sc, keep = make_scene(scene)
        out = np.zeros(len(e), dtype=np.uint8)
        if threads > 1:
            lib().or_validate_edges_mt(C.byref(self.m), C.byref(sc), _ptr(e), len(e), g, _ptr(out), threads)
        else:
            lib().or_validate_edges(C.byref(self.m), C.byref(sc), _ptr(e), len(e), g, _ptr(out))
        return out.astype(bool)

    def plan(self, scene, start, goals, settings: dict, mode=0, path_cap=4096):
        st = make_settings(settings)
        sc, keep = make_scene(scene)
        start = np.ascontiguousarray(start, dtype=np.float32)
        goals = np.ascontiguousarray(np.asarray(goals, dtype=np.float32).reshape(-1, self.dof))
        path = np.zeros((path_cap, self.dof), dtype=np.float32)
        res = Result()
        rc = lib().or_plan(C.byref(self.m), C.byref(sc), _ptr(start), _ptr(goals), len(goals), C.byref(st), mode,
                           _ptr(path), path_cap, C.byref(res))
        return rc, res, path[:res.path_nodes].copy()

    def rrt_connect(self, scene, start, goal, granularity, rng, max_samples, seed, path_cap=8192):
        sc, keep = make_scene(scene)
        start = np.ascontiguousarray(start, dtype=np.float32)
        goal = np.ascontiguousarray(goal, dtype=np.float32)
        path = np.zeros((path_cap, self.dof), dtype=np.float32)
        res = Result()
        rc = lib().or_rrt_connect(C.byref(self.m), C.byref(sc), _ptr(start), _ptr(goal), granularity, rng,
                                  max_samples, seed, _ptr(path), path_cap, C.byref(res))
        return rc, res, path[:res.path_nodes].copy()

    def revalidate(self, scene, path, g):
        sc, keep = make_scene(scene)
        p = np.ascontiguousarray(path, dtype=np.float32)
        return lib().or_revalidate_path(C.byref(self.m), C.byref(sc), _ptr(p), len(p), g)


def make_scene(scene):
    sc = Scene()
    keep = []
    for name, w in (("spheres", 4), ("capsules", 8), ("z_capsules", 8), ("cuboids", 15), ("z_cuboids", 15)):
        a = np.ascontiguousarray(np.asarray(getattr(scene, name), dtype=np.float32).reshape(-1, w))
        keep.append(a)
        setattr(sc, name, a.ctypes.data_as(_f32p) if len(a) else None)
        setattr(sc, "n_" + name, len(a))
    sc._keep = keep
    return sc, keep


def make_settings(d: dict) -> Settings:
    st = Settings()
    st.max_samples = d.get("max_samples", 1000000)
    st.max_iters = d.get("max_iters", 1000000)
    st.num_new_configs = d.get("num_new_configs", 600)
    st.granularity = d.get("granularity", 64)
    st.range = d.get("range", 0.5)
    st.balance = d.get("balance", 1)
    st.tree_ratio = d.get("tree_ratio", 0.5)
    st.dynamic_domain = int(d.get("dynamic_domain", True))
    st.dd_alpha = d.get("dd_alpha", 0.0001)
    st.dd_radius = d.get("dd_radius", 4.0)
    st.dd_min_radius = d.get("dd_min_radius", 1.0)
    st.seed = d.get("seed", 1)
    return st


def set_fused(on: bool) -> None:
    """Sphere tests rounded like the HIP kernels (True) or the reference (False)."""
    lib().or_set_fused(1 if on else 0)


# Planted planner mistakes, in the order of the oracle's PLANT_* list (tests
# show that the known answers catch each one).
PLANTS = ("iter1_tree", "b1_ratio", "b1_split", "b2_rule", "dd_sign", "dd_skip", "chain_drop", "steer",
          "path_order", "n_ext")


def set_plant(name: str | None) -> None:
    """Make or_plan commit a planted mistake (tests only); None restores it."""
    k = 0 if name is None else PLANTS.index(name) + 1
    assert lib().or_set_plant(k) == len(PLANTS) + 1


def halton(dof: int, seed: int, slot: int, n: int) -> np.ndarray:
    state = (C.c_float * (3 * MAXD))()
    lib().or_halton_init(C.byref(state), dof, seed, slot)
    out = np.zeros((n, dof), dtype=np.float32)
    row = (C.c_float * MAXD)()
    for i in range(n):
        lib().or_halton_next(C.byref(state), dof, C.byref(row))
        out[i] = np.frombuffer(row, dtype=np.float32)[:dof]
    return out
