"""ctypes binding of libprrtc.so (the C ABI in include/prrtc.h).

The library must have been built (``python -m prrtc_amd.build``); loading
fails loudly otherwise -- there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .robots import DOF, ROBOT_IDS
from .scenes import Scene

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprrtc.so")

PRRTC_OK = 0
ERRORS = {-1: "EINVAL", -2: "EDEVICE", -3: "ENOMEM", -4: "EOVERFLOW", -5: "EBUSY", -6: "EAGAIN"}
PRRTC_EBUSY = -5
PRRTC_EAGAIN = -6
PRRTC_ETIMEDOUT = -7
PRRTC_ECANCELED = -8
MAX_DOF = 14

_f32p = C.POINTER(C.c_float)


class CScene(C.Structure):
    _fields_ = [
        ("spheres", _f32p), ("n_spheres", C.c_int32),
        ("capsules", _f32p), ("n_capsules", C.c_int32),
        ("z_capsules", _f32p), ("n_z_capsules", C.c_int32),
        ("cuboids", _f32p), ("n_cuboids", C.c_int32),
        ("z_cuboids", _f32p), ("n_z_cuboids", C.c_int32),
    ]


class CSettings(C.Structure):
    _fields_ = [
        ("max_samples", C.c_int32), ("max_iters", C.c_int32), ("num_new_configs", C.c_int32),
        ("granularity", C.c_int32), ("range", C.c_float), ("balance", C.c_int32), ("tree_ratio", C.c_float),
        ("dynamic_domain", C.c_int32), ("dd_alpha", C.c_float), ("dd_radius", C.c_float),
        ("dd_min_radius", C.c_float), ("seed", C.c_uint64),
    ]


class CResult(C.Structure):
    _fields_ = [
        ("solved", C.c_int32), ("start_tree_size", C.c_int32), ("goal_tree_size", C.c_int32),
        ("path_length", C.c_int32), ("iters", C.c_int32), ("cost", C.c_float), ("path_nodes", C.c_int32),
        ("status", C.c_int32), ("wall_ns", C.c_int64), ("kernel_ns", C.c_int64), ("copy_ns", C.c_int64),
        ("edge_checks", C.c_int64), ("pose_checks", C.c_int64),
    ]


class CProblem(C.Structure):
    _fields_ = [
        ("robot", C.c_int32), ("start", _f32p), ("goals", _f32p), ("n_goals", C.c_int32), ("scene", CScene),
    ]


class CBatchStats(C.Structure):
    _fields_ = [
        ("wall_ns", C.c_int64), ("kernel_ns", C.c_int64), ("edge_checks", C.c_int64), ("solved", C.c_int32),
        ("rounds", C.c_int32), ("pose_checks", C.c_int64),
    ]


class CTuning(C.Structure):
    _fields_ = [("waves_per_cu", C.c_int32), ("waves_per_task", C.c_int32), ("batch_waves_per_task", C.c_int32),
                ("crowd_above", C.c_int32)]


class CServerConfig(C.Structure):
    _fields_ = [
        ("budget_bytes", C.c_uint64), ("max_resident", C.c_int32), ("max_pending", C.c_int32),
        ("slice_us", C.c_int32), ("path_cap", C.c_int32), ("young_rounds", C.c_int32), ("deadline_ms", C.c_int32),
    ]


class CServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("solved", C.c_int64), ("pending", C.c_int32),
        ("resident", C.c_int32), ("resident_bytes", C.c_uint64), ("peak_resident_bytes", C.c_uint64),
        ("slices", C.c_int64), ("kernel_ns", C.c_int64), ("edge_checks", C.c_int64), ("timed_out", C.c_int64),
        ("cancelled", C.c_int64),
    ]


class CRequestTiming(C.Structure):
    _fields_ = [("queued_ns", C.c_int64), ("resident_ns", C.c_int64), ("slices", C.c_int32)]


# (name, restype, argtypes) of every symbol include/prrtc.h declares
SIGNATURES = [
    ("prrtc_settings_default", None, [C.POINTER(CSettings)]),
    ("prrtc_version", C.c_int, []),
    ("prrtc_last_error", C.c_char_p, []),
    ("prrtc_robot_dof", C.c_int, [C.c_int]),
    ("prrtc_robot_num_spheres", C.c_int, [C.c_int]),
    ("prrtc_scene_packed_floats", C.c_size_t, [C.POINTER(CScene)]),
    ("prrtc_scene_pack", C.c_int, [C.POINTER(CScene), C.c_void_p, C.c_size_t]),
    ("prrtc_fk_spheres", C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    ("prrtc_check_poses", C.c_int, [C.c_int, C.POINTER(CScene), C.c_void_p, C.c_int, C.c_void_p]),
    ("prrtc_validate_edges", C.c_int, [C.c_int, C.POINTER(CScene), C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    ("prrtc_validate_edges_async", C.c_int,
     [C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    ("prrtc_nearest_async", C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_void_p]),
    ("prrtc_solve", C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(CScene), C.POINTER(CSettings),
                              C.POINTER(CResult), C.c_void_p, C.c_int]),
    ("prrtc_solve_batch", C.c_int, [C.c_int, C.POINTER(CProblem), C.POINTER(CSettings), C.c_int,
                                    C.POINTER(CResult), C.c_void_p, C.c_int, C.POINTER(CBatchStats)]),
    ("prrtc_solve_batch_rounds", C.c_int, [C.c_int, C.POINTER(CProblem), C.POINTER(CSettings), C.c_int,
                                           C.POINTER(CResult), C.c_void_p, C.c_int, C.POINTER(CBatchStats)]),
    ("prrtc_set_tuning", C.c_int, [C.c_int, C.POINTER(CTuning)]),
    ("prrtc_get_tuning", C.c_int, [C.c_int, C.POINTER(CTuning)]),
    ("prrtc_server_config_default", None, [C.POINTER(CServerConfig)]),
    ("prrtc_server_create", C.c_int, [C.POINTER(CServerConfig), C.POINTER(C.c_void_p)]),
    ("prrtc_server_submit", C.c_int, [C.c_void_p, C.POINTER(CProblem), C.POINTER(CSettings),
                                      C.POINTER(C.c_uint64)]),
    ("prrtc_server_poll", C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.POINTER(CResult), C.c_void_p, C.c_int32,
                                    C.POINTER(CRequestTiming)]),
    ("prrtc_server_get_stats", C.c_int, [C.c_void_p, C.POINTER(CServerStats)]),
    ("prrtc_server_cancel", C.c_int, [C.c_void_p, C.c_uint64]),
    ("prrtc_solve_noalloc", C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(CScene),
                                      C.POINTER(CSettings), C.POINTER(CResult), C.c_void_p, C.c_int]),
    ("prrtc_server_destroy", C.c_int, [C.c_void_p]),
]

_lib = None


class PrrtcError(RuntimeError):
    """A call returned a PRRTC_E* status; `code` holds it."""

    def __init__(self, msg: str, code: int | None = None):
        super().__init__(msg)
        self.code = code


def load():
    """Load libprrtc.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m prrtc_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != PRRTC_OK:
        msg = load().prrtc_last_error().decode()
        raise PrrtcError(f"{what} failed ({ERRORS.get(rc, rc)}): {msg}", rc)


def robot_id(robot) -> int:
    return ROBOT_IDS[robot] if isinstance(robot, str) else int(robot)


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def c_scene(scene: Scene):
    """A CScene pointing into the scene's arrays (the arrays are kept alive)."""
    cs = CScene()
    keep = []
    for name, w in (("spheres", 4), ("capsules", 8), ("z_capsules", 8), ("cuboids", 15), ("z_cuboids", 15)):
        a = _f(getattr(scene, name)).reshape(-1, w)
        keep.append(a)
        setattr(cs, name, a.ctypes.data_as(_f32p) if len(a) else None)
        setattr(cs, "n_" + name, len(a))
    cs._keep = keep
    return cs


def c_settings(d: dict) -> CSettings:
    st = CSettings()
    load().prrtc_settings_default(C.byref(st))
    for k, v in d.items():
        if not hasattr(st, k):
            raise KeyError(f"unknown planner setting {k!r}")
        setattr(st, k, int(v) if k in ("max_samples", "max_iters", "num_new_configs", "granularity", "balance",
                                        "dynamic_domain", "seed") else float(v))
    return st


def settings_default() -> dict:
    st = CSettings()
    load().prrtc_settings_default(C.byref(st))
    return {name: getattr(st, name) for name, _ in CSettings._fields_}


def set_tuning(robot, waves_per_cu: int = 0, waves_per_task: int = 0, batch_waves_per_task: int = 0,
               crowd_above: int = 0) -> None:
    """Launch shapes of the persistent planner for `robot` (every robot when
    robot is None); 0 = default.  See prrtc_tuning in include/prrtc.h."""
    t = CTuning(int(waves_per_cu), int(waves_per_task), int(batch_waves_per_task), int(crowd_above))
    _check(load().prrtc_set_tuning(-1 if robot is None else robot_id(robot), C.byref(t)), "prrtc_set_tuning")


def get_tuning(robot) -> dict:
    t = CTuning()
    _check(load().prrtc_get_tuning(robot_id(robot), C.byref(t)), "prrtc_get_tuning")
    return {name: getattr(t, name) for name, _ in CTuning._fields_}


def pack_scene(scene: Scene) -> np.ndarray:
    cs = c_scene(scene)
    n = load().prrtc_scene_packed_floats(C.byref(cs))
    out = np.zeros(n, dtype=np.float32)
    _check(load().prrtc_scene_pack(C.byref(cs), out.ctypes.data, n), "prrtc_scene_pack")
    return out


def fk_spheres(robot, poses) -> np.ndarray:
    rid = robot_id(robot)
    q = _f(poses).reshape(-1, load().prrtc_robot_dof(rid))
    S = load().prrtc_robot_num_spheres(rid)
    out = np.zeros((len(q), S, 3), dtype=np.float32)
    _check(load().prrtc_fk_spheres(rid, q.ctypes.data, len(q), out.ctypes.data), "prrtc_fk_spheres")
    return out


def check_poses(robot, scene: Scene, poses) -> np.ndarray:
    rid = robot_id(robot)
    q = _f(poses).reshape(-1, load().prrtc_robot_dof(rid))
    out = np.zeros(len(q), dtype=np.uint8)
    cs = c_scene(scene)
    _check(load().prrtc_check_poses(rid, C.byref(cs), q.ctypes.data, len(q), out.ctypes.data), "prrtc_check_poses")
    return out.astype(bool)


def validate_edges(robot, scene: Scene, edges, granularity: int) -> np.ndarray:
    rid = robot_id(robot)
    D = load().prrtc_robot_dof(rid)
    e = _f(edges).reshape(-1, 2, D)
    out = np.zeros(len(e), dtype=np.uint8)
    cs = c_scene(scene)
    _check(load().prrtc_validate_edges(rid, C.byref(cs), e.ctypes.data, len(e), int(granularity), out.ctypes.data),
           "prrtc_validate_edges")
    return out.astype(bool)


def validate_edges_async(robot, d_scene: int, scene_floats: int, d_edges: int, n: int, granularity: int,
                         d_out: int, stream: int = 0) -> None:
    """Device-pointer form: all pointers are device addresses (ints)."""
    _check(load().prrtc_validate_edges_async(robot_id(robot), d_scene, scene_floats, d_edges, n, granularity,
                                             d_out, stream), "prrtc_validate_edges_async")


def nearest_async(dof: int, d_trees: int, cap: int, n_nodes: int, n_trees: int, d_queries: int,
                  queries_per_tree: int, waves: int, d_idx: int, d_dist2: int, stream: int = 0) -> None:
    """Nearest-node search on device buffers (all pointers device addresses)."""
    _check(load().prrtc_nearest_async(dof, d_trees, cap, n_nodes, n_trees, d_queries, queries_per_tree, waves,
                                      d_idx, d_dist2, stream), "prrtc_nearest_async")


def _result_dict(r: CResult) -> dict:
    return {name: getattr(r, name) for name, _ in CResult._fields_}


def solve(robot, start, goals, scene: Scene, settings: dict | None = None, path_cap: int = 8192):
    """pRRTC::solve<Robot> (src/planning/Planners.hh:75-78) through the C ABI.

    Returns (result dict, path [rows, dof] goal -> start)."""
    rid = robot_id(robot)
    D = load().prrtc_robot_dof(rid)
    start = _f(start).reshape(D)
    goals = _f(goals).reshape(-1, D)
    st = c_settings(settings or {})
    cs = c_scene(scene)
    res = CResult()
    path = np.zeros((path_cap, D), dtype=np.float32)
    rc = load().prrtc_solve(rid, start.ctypes.data, goals.ctypes.data, len(goals), C.byref(cs), C.byref(st),
                            C.byref(res), path.ctypes.data, path_cap)
    _check(rc, "prrtc_solve")
    return _result_dict(res), path[:res.path_nodes].copy()


def solve_noalloc(robot, start, goals, scene: Scene, settings: dict | None = None, path_cap: int = 8192):
    """prrtc_solve_noalloc: solve() with its device state in the library's
    static scratch array (no device or pinned allocation)."""
    rid = robot_id(robot)
    D = load().prrtc_robot_dof(rid)
    start = _f(start).reshape(D)
    goals = _f(goals).reshape(-1, D)
    st = c_settings(settings or {})
    cs = c_scene(scene)
    res = CResult()
    path = np.zeros((path_cap, D), dtype=np.float32)
    rc = load().prrtc_solve_noalloc(rid, start.ctypes.data, goals.ctypes.data, len(goals), C.byref(cs), C.byref(st),
                                    C.byref(res), path.ctypes.data, path_cap)
    _check(rc, "prrtc_solve_noalloc")
    return _result_dict(res), path[:res.path_nodes].copy()


def solve_batch(problems, settings, path_cap: int = 4096, graph_rounds: bool = False):
    """Solve many problems in one persistent launch.

    problems: list of dicts {robot, start, goals, scene}; settings: a dict or
    a list of dicts (one per problem).  Returns (results, paths, stats).
    graph_rounds: run the rounds as graph-replayed grids instead
    (prrtc_solve_batch_rounds, the non-persistent comparator)."""
    n = len(problems)
    cps = (CProblem * n)()
    keep = []
    dofs = []
    for i, p in enumerate(problems):
        rid = robot_id(p["robot"])
        D = load().prrtc_robot_dof(rid)
        dofs.append(D)
        s = _f(p["start"]).reshape(D)
        g = _f(p["goals"]).reshape(-1, D)
        cs = c_scene(p["scene"])
        keep += [s, g, cs]
        cps[i].robot = rid
        cps[i].start = s.ctypes.data_as(_f32p)
        cps[i].goals = g.ctypes.data_as(_f32p)
        cps[i].n_goals = len(g)
        cps[i].scene = cs
    if isinstance(settings, dict):
        sts = (CSettings * 1)(c_settings(settings))
        n_st = 1
    else:
        sts = (CSettings * n)(*[c_settings(s) for s in settings])
        n_st = n
    res = (CResult * n)()
    paths = np.zeros((n, path_cap * MAX_DOF), dtype=np.float32)
    stats = CBatchStats()
    fn = "prrtc_solve_batch_rounds" if graph_rounds else "prrtc_solve_batch"
    rc = getattr(load(), fn)(n, cps, sts, n_st, res, paths.ctypes.data, path_cap, C.byref(stats))
    _check(rc, fn)
    results = [_result_dict(res[i]) for i in range(n)]
    # block i holds path_nodes rows of dof floats, packed
    out_paths = [paths[i, :res[i].path_nodes * dofs[i]].reshape(-1, dofs[i]).copy() for i in range(n)]
    st = {name: getattr(stats, name) for name, _ in CBatchStats._fields_}
    return results, out_paths, st


def _c_problem(p, keep):
    rid = robot_id(p["robot"])
    D = load().prrtc_robot_dof(rid)
    s = _f(p["start"]).reshape(D)
    g = _f(p["goals"]).reshape(-1, D)
    cs = c_scene(p["scene"])
    keep += [s, g, cs]
    cp = CProblem()
    cp.robot = rid
    cp.start = s.ctypes.data_as(_f32p)
    cp.goals = g.ctypes.data_as(_f32p)
    cp.n_goals = len(g)
    cp.scene = cs
    return cp, D


class Server:
    """Continuous-batching server (prrtc_server_* in include/prrtc.h).

    submit() copies the problem and returns a ticket; poll(ticket) returns
    (result, path, timing) once the problem has stopped, else None."""

    def __init__(self, budget_bytes=0, max_resident=0, max_pending=0, slice_us=0, path_cap=0, young_rounds=0,
                 deadline_ms=0):
        cfg = CServerConfig(budget_bytes, max_resident, max_pending, slice_us, path_cap, young_rounds, deadline_ms)
        h = C.c_void_p()
        _check(load().prrtc_server_create(C.byref(cfg), C.byref(h)), "prrtc_server_create")
        self._h = h
        self._dof = {}
        self.path_cap = path_cap or 1024

    def submit(self, problem: dict, settings: dict | None = None) -> int:
        keep = []
        cp, D = _c_problem(problem, keep)
        st = c_settings(settings or {})
        t = C.c_uint64()
        _check(load().prrtc_server_submit(self._h, C.byref(cp), C.byref(st), C.byref(t)), "prrtc_server_submit")
        self._dof[t.value] = D
        return t.value

    def poll(self, ticket: int, wait_ms: int = -1):
        D = self._dof[ticket]
        res = CResult()
        tm = CRequestTiming()
        path = np.zeros((self.path_cap, D), dtype=np.float32)
        rc = load().prrtc_server_poll(self._h, ticket, wait_ms, C.byref(res), path.ctypes.data, self.path_cap,
                                      C.byref(tm))
        if rc == PRRTC_EAGAIN:
            return None
        _check(rc, "prrtc_server_poll")
        del self._dof[ticket]
        timing = {name: getattr(tm, name) for name, _ in CRequestTiming._fields_}
        return _result_dict(res), path[:res.path_nodes].copy(), timing

    def cancel(self, ticket: int) -> None:
        """End the request unsolved (result status PRRTC_ECANCELED) unless it
        has finished already; poll it as usual."""
        _check(load().prrtc_server_cancel(self._h, ticket), "prrtc_server_cancel")

    def stats(self) -> dict:
        st = CServerStats()
        _check(load().prrtc_server_get_stats(self._h, C.byref(st)), "prrtc_server_get_stats")
        return {name: getattr(st, name) for name, _ in CServerStats._fields_}

    def close(self):
        if self._h:
            load().prrtc_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
