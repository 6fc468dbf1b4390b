"""Robot model tables (joint tree, collision spheres, check structure).

``<robot>.json`` for Panda, Fetch and Baxter is produced once by ``fit_model``
from the reference's generated FK code; see that module for the derivation.
Robot ids 0-3 follow the reference's explicit instantiations
(src/planning/pRRTC.cu:768-771).  UR5 (id 4) is not in the reference: its
table is compiled by ``urdf_compile`` from ``ur5.urdf`` and
``ur5_spheres.json``.
"""
from __future__ import annotations

import functools
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

ROBOT_IDS = {"sphere": 0, "panda": 1, "fetch": 2, "baxter": 3, "ur5": 4}
DOF = {"sphere": 3, "panda": 7, "fetch": 8, "baxter": 14, "ur5": 6}

# The 3-DoF point robot: one sphere of radius 1 at q (Robots.hh:190-237,
# utils.cuh:330-334); joint limits [-5, 5]^3.
SPHERE_ROBOT = {"name": "sphere", "dof": 3, "lower": [-5.0, -5.0, -5.0], "range": [10.0, 10.0, 10.0],
                "radius": 1.0}


@functools.lru_cache(maxsize=None)
def load_model(name: str) -> dict:
    if name == "sphere":
        return dict(SPHERE_ROBOT)
    with open(os.path.join(HERE, name + ".json")) as fh:
        return json.load(fh)


def limits(name: str):
    m = load_model(name)
    return np.array(m["lower"], dtype=np.float32), np.array(m["range"], dtype=np.float32)


def fixed_spheres(name: str) -> np.ndarray:
    """Spheres that never move (frame -1): the robot's base, [n, 4]."""
    if name == "sphere":
        return np.zeros((0, 4), dtype=np.float32)
    m = load_model(name)
    rows = [s["local"] + [s["r"]] for s in m["spheres"] if s["frame"] == -1]
    return np.array(rows, dtype=np.float32).reshape(-1, 4)
