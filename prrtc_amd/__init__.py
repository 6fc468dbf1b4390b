"""prrtc_amd: pRRTC's planning hot path, MI355X-native (gfx950).

Host API over the C ABI of ``libprrtc.so`` (include/prrtc.h):

    fk_spheres, check_poses, validate_edges, validate_edges_async,
    solve (pRRTC::solve<Robot>), solve_batch, Server (continuous batching),
    pack_scene

plus the robot model tables (``prrtc_amd.robots``) and seeded synthetic
scenes (``prrtc_amd.scenes``).
"""
from .lib import (PrrtcError, Server, check_poses, fk_spheres, get_tuning, load, nearest_async, pack_scene, set_tuning,
                  settings_default, solve, solve_batch, solve_noalloc, validate_edges, validate_edges_async)
from .robots import DOF, ROBOT_IDS, limits, load_model
from .scenes import Scene

__all__ = [
    "PrrtcError", "Server", "check_poses", "fk_spheres", "get_tuning", "set_tuning", "load", "nearest_async", "pack_scene", "settings_default", "solve", "solve_batch", "solve_noalloc",
    "validate_edges", "validate_edges_async", "DOF", "ROBOT_IDS", "limits", "load_model", "Scene",
]
