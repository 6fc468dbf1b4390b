import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
ROBOTS = ("panda", "fetch", "baxter")
SCENE_KEYS = ("spheres", "capsules", "z_capsules", "cuboids", "z_cuboids")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def golden_scene(npz, prefix):
    from prrtc_amd.scenes import Scene
    return Scene.from_arrays({k: npz[f"{prefix}__{k}"] for k in SCENE_KEYS})


NEAR = 1e-6   # poses closer than this to contact may flip verdicts (float rounding)


def edge_poses(edges, g):
    """Poses a + (k+1)(b-a)/g, k = 0..g-1, of each edge (pRRTC.cu:438-439), float32."""
    a, b = edges[:, 0], edges[:, 1]
    delta = (b - a) / np.float32(g)
    k = np.arange(1, g + 1, dtype=np.float32)[None, :, None]
    return (a[:, None, :] + k * delta[:, None, :]).astype(np.float32)


def edge_margins(R, scene, edges, g):
    """Smallest distance to contact (m) over the poses of each edge, from the
    oracle's exhaustive clearance (every sphere against every obstacle and
    self pair): an edge verdict may only flip where this is below NEAR."""
    if len(edges) == 0:
        return np.zeros(0, np.float32)
    q = edge_poses(np.asarray(edges, np.float32), g)
    _, m = R.check(scene, q.reshape(-1, q.shape[-1]), margin=True)
    return m.reshape(len(edges), g).min(axis=1)


@pytest.fixture(scope="session")
def oracle_robots():
    from oracle import oracle
    from prrtc_amd.robots import load_model
    oracle.build()
    return {r: oracle.OracleRobot(load_model(r)) for r in ("sphere",) + ROBOTS + ("ur5",)}


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, loaded; skips nothing -- a missing library is an error.

    torch (used by the device-buffer tests) brings its own HIP runtime, which
    must see the GPU before the library's runtime initialises it."""
    import torch
    if torch.cuda.is_available():
        torch.zeros(1, device="cuda:0")
    import prrtc_amd
    prrtc_amd.load()
    return prrtc_amd
