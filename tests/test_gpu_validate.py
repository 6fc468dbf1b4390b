"""GPU parity of FK, pose checks and edge validation (through the C ABI)
against the reference's golden vectors and the CPU oracle."""
import zlib

import numpy as np
import pytest

from conftest import NEAR, ROBOTS, edge_margins, golden, golden_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("robot", ROBOTS)
def test_fk_centres(robot, gpu):
    g = golden(f"fk_{robot}.npz")
    c = gpu.fk_spheres(robot, g["q"])
    err = float(np.abs(c - g["centres"]).max())
    print(f"{robot}: max FK centre error vs reference {err:.3e} m")
    assert err < 1e-5


@pytest.mark.parametrize("robot", ROBOTS)
@pytest.mark.parametrize("scene", ["spheres12", "obb_cyl", "shelf", "ztables"])
def test_pose_verdicts(robot, scene, gpu, oracle_robots):
    g = golden(f"verdict_{robot}.npz")
    sc = golden_scene(g, scene)
    free = gpu.check_poses(robot, sc, g["poses"])
    _, margin = oracle_robots[robot].check(sc, g["poses"], margin=True)
    mism = free != g[f"{scene}__free"]
    print(f"{robot}/{scene}: {mism.sum()} verdict flips, {(mism & (margin < NEAR)).sum()} of them near contact")
    assert not np.any(mism & (margin >= NEAR))


@pytest.mark.parametrize("robot", ROBOTS)
@pytest.mark.parametrize("g", [32, 64, 128])
def test_edge_verdicts_golden(robot, g, gpu):
    e = golden(f"edge_{robot}.npz")
    sc = golden_scene(e, "scene")
    free = gpu.validate_edges(robot, sc, e["edges"], g)
    assert (free != e[f"free_g{g}"]).sum() == 0


@pytest.mark.parametrize("robot", ("sphere",) + ROBOTS)
@pytest.mark.parametrize("g", [1, 7, 32, 64, 100, 128, 256])
def test_edge_verdicts_oracle(robot, g, gpu, oracle_robots):
    """Random edges (ragged granularities included) against the oracle."""
    from prrtc_amd import scenes as S
    from prrtc_amd.robots import fixed_spheres, limits
    lo, span = limits(robot) if robot != "sphere" else (np.full(3, -5, np.float32), np.full(3, 10, np.float32))
    rng = np.random.default_rng(zlib.crc32(f"{robot}/{g}".encode()))
    if robot == "sphere":
        # the point robot lives in [-5, 5]^3: obstacles sized for that box
        sc = S.Scene.build(
            spheres=[S.sphere(rng.uniform(-4, 4, 3), rng.uniform(0.5, 1.5)) for _ in range(6)],
            cuboids=[S.cuboid(rng.uniform(-4, 4, 3), rng.uniform(-1, 1, 3), rng.uniform(0.3, 1.2, 3))
                     for _ in range(4)],
            capsules=[S.capsule_endpoints(rng.uniform(-4, 4, 3), rng.uniform(-4, 4, 3), 0.4) for _ in range(3)])
    else:
        sc = S.scene_mixed(g, 8, 32, avoid=fixed_spheres(robot))
    R = oracle_robots[robot]
    # edges from free starting poses so that both outcomes are common
    cand = (lo + span * rng.random((4096, len(lo)), dtype=np.float32)).astype(np.float32)
    cand = cand[R.check(sc, cand)][:300]
    d = rng.normal(size=cand.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    b = np.clip(cand + d * rng.uniform(0.05, 0.6, size=(len(cand), 1)), lo, lo + span).astype(np.float32)
    edges = np.stack([cand, b], 1)
    want = R.validate_edges(sc, edges, g)
    got = gpu.validate_edges(robot, sc, edges, g)
    assert 0 < want.sum() < len(want) or len(want) < 20
    # a verdict may differ only on an edge with a pose within NEAR of contact
    flips = np.nonzero(got != want)[0]
    m = edge_margins(R, sc, edges[flips], g)
    print(f"{robot}/g{g}: {len(flips)} edge flips of {len(want)}, all within {NEAR} m of contact: "
          f"{bool(np.all(m < NEAR))}")
    assert np.all(m < NEAR), f"edges {flips[m >= NEAR].tolist()} differ {m[m >= NEAR]} m from contact"


def test_empty_inputs(gpu):
    from prrtc_amd import scenes as S
    sc = S.scene_spheres(0)
    assert gpu.validate_edges("panda", sc, np.zeros((0, 2, 7), np.float32), 32).shape == (0,)
    assert gpu.check_poses("panda", sc, np.zeros((0, 7), np.float32)).shape == (0,)
    # an empty scene: only self-collision decides
    q = np.zeros((4, 7), np.float32)
    assert gpu.check_poses("panda", S.Scene(), q).shape == (4,)


def test_nonfinite_rejected(gpu):
    from prrtc_amd import scenes as S
    e = np.zeros((2, 2, 7), np.float32)
    e[1, 1, 3] = np.nan
    with pytest.raises(gpu.PrrtcError):
        gpu.validate_edges("panda", S.scene_spheres(0), e, 32)


def test_cull_boundary(gpu, oracle_robots):
    """Poses a fraction of a millimetre to 2 mm from contact where the
    bounding balls of the culls are tight: beyond a cuboid's corners along
    its diagonals and beyond a capsule's ends along its axis.  The culled
    kernel must give the oracle's verdict on every one (none is within
    1e-6 m of contact)."""
    from prrtc_amd import scenes as S
    rng = np.random.default_rng(7)
    c = np.array([0.3, -0.2, 0.5], np.float32)
    h = np.array([0.4, 0.7, 0.25], np.float32)
    cub = S.cuboid(c, (0.3, -0.5, 1.1), h)
    R = cub[3:12].reshape(3, 3).T          # columns are the box axes
    p1, p2 = np.array([-2.0, 1.5, 0.2], np.float32), np.array([-1.0, 2.5, 1.4], np.float32)
    cap = S.capsule_endpoints(p1, p2, 0.3)
    sc = S.Scene.build(cuboids=[cub], capsules=[cap])
    deltas = np.array([-2e-3, -5e-4, -1e-4, 1e-4, 5e-4, 2e-3], np.float32)
    pts = []
    for sx in (-1, 1):
        for sy in (-1, 1):
            for sz in (-1, 1):
                corner = c + R @ (h * np.array([sx, sy, sz], np.float32))
                u = R @ (h * np.array([sx, sy, sz], np.float32))
                u /= np.linalg.norm(u)
                for d in deltas:
                    pts.append(corner + u * (1.0 + d))                      # robot radius 1
                    w = u + 0.05 * rng.normal(size=3)
                    pts.append(corner + w / np.linalg.norm(w) * (1.0 + d))
    axis = (p2 - p1) / np.linalg.norm(p2 - p1)
    for end, s in ((p1, -1.0), (p2, 1.0)):
        for d in deltas:
            pts.append(end + s * axis * (1.3 + d))                            # robot 1 + capsule 0.3
    poses = np.array(pts, np.float32)
    R_or = oracle_robots["sphere"]
    want, margin = R_or.check(sc, poses, margin=True)
    got = gpu.check_poses("sphere", sc, poses)
    assert np.all(margin >= 1e-6)
    assert 0 < want.sum() < len(want)
    assert np.array_equal(got, want), f"{(got != want).sum()} of {len(want)} verdicts differ"
