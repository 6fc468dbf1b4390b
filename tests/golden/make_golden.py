"""Write the golden vectors under tests/golden/ from the reference's own FK.

Run once, in a container that holds the reference sources:

    python tests/golden/make_golden.py /root/reference

The reference's forward kinematics and collision checks exist only as
generated straight-line code (src/planning/utils.cuh:336-33131).  The
reference cannot be compiled here (it needs the CUDA headers and cuRAND), so
this script reads that code as text with ``prrtc_amd.robots.fkcc_text`` and
evaluates its arithmetic with the C++ types the code declares (binary32 for
``float`` values, binary64 where ``auto`` meets a double literal), then
applies the reference's check semantics:

* ``sphere_environment_in_collision`` (utils.cuh:211-242): OR over spheres,
  capsules, z-aligned capsules, cuboids, z-aligned cuboids, each test in
  squared form, ``< 0`` meaning collision, evaluated in binary32 with the
  formulas of utils.cuh:58-208;
* the hierarchy of each generated check (gate sphere, then leaf spheres only
  when the gate hits; self-collision gate pair, then sphere pairs);
* edges: g poses ``a + (k+1) * ((b - a) / g)``, k = 0..g-1 (pRRTC.cu:430-440),
  an edge is free iff every pose is free.

Outputs (numpy ``.npz``, no pickles):

* ``fk_<robot>.npz``      q [256, D], centres [256, S, 3], radii [S]
  (sphere order = the model table's ``spheres``);
* ``verdict_<robot>.npz`` per scene: the scene tables, poses [2048, D] and
  free [2048] (bool);
* ``edge_<robot>.npz``    scene tables, edges [192, 2, D], free flags at
  g = 32, 64, 128.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from prrtc_amd import scenes as S  # noqa: E402
from prrtc_amd.robots import fkcc_text as F  # noqa: E402
from prrtc_amd.robots import fixed_spheres  # noqa: E402
from prrtc_amd.robots.fit_model import read_limits  # noqa: E402

f32 = np.float32


# ---------------------------------------------------------------------------
# reference check semantics in binary32 (utils.cuh:19-242)

def _sql2(ax, ay, az, bx, by, bz):
    xs, ys, zs = f32(ax - bx), f32(ay - by), f32(az - bz)
    return f32(f32(f32(xs * xs) + f32(ys * ys)) + f32(zs * zs))


def env_hit(scene: S.Scene, x, y, z, r):
    """sphere_environment_in_collision on arrays of sphere centres."""
    x, y, z = (np.asarray(v, dtype=f32) for v in (x, y, z))
    r = f32(r)
    rsq = f32(r * r)
    hit = np.zeros(x.shape, dtype=bool)
    for s in scene.spheres:
        d = _sql2(s[0], s[1], s[2], x, y, z)
        rs = f32(s[3] + r)
        hit |= f32(d - f32(rs * rs)) < 0
    for caps, zal in ((scene.capsules, False), (scene.z_capsules, True)):
        for c in caps:
            x1, y1, z1, xv, yv, zv, cr, rdv = (f32(v) for v in c)
            if zal:
                dot = f32(f32(z - z1) * zv)
                cdf = np.clip(f32(dot * rdv), f32(0), f32(1)).astype(f32)
                d = _sql2(x, y, z, x1, y1, f32(z1 + f32(zv * cdf)))
            else:
                dot = f32(f32(f32(f32(x - x1) * xv) + f32(f32(y - y1) * yv)) + f32(f32(z - z1) * zv))
                cdf = np.clip(f32(dot * rdv), f32(0), f32(1)).astype(f32)
                d = _sql2(x, y, z, f32(x1 + f32(xv * cdf)), f32(y1 + f32(yv * cdf)), f32(z1 + f32(zv * cdf)))
            rs = f32(r + cr)
            hit |= f32(d - f32(rs * rs)) < 0
    for cubs, zal in ((scene.cuboids, False), (scene.z_cuboids, True)):
        for c in cubs:
            c = c.astype(f32)
            xs, ys, zs = f32(x - c[0]), f32(y - c[1]), f32(z - c[2])
            acc = []
            for a in range(3):
                ax, ay, az, hr = c[3 + 3 * a], c[4 + 3 * a], c[5 + 3 * a], c[12 + a]
                if zal and a == 2:
                    proj = np.abs(zs)
                elif zal:
                    proj = np.abs(f32(f32(ax * xs) + f32(ay * ys)))
                else:
                    proj = np.abs(f32(f32(f32(ax * xs) + f32(ay * ys)) + f32(az * zs)))
                acc.append(np.maximum(f32(0), f32(proj - hr)).astype(f32))
            ssum = f32(f32(f32(acc[0] * acc[0]) + f32(acc[1] * acc[1])) + f32(acc[2] * acc[2]))
            hit |= f32(ssum - rsq) < 0
    return hit


def self_hit(a, b, ra, rb):
    d = _sql2(a[..., 0], a[..., 1], a[..., 2], b[..., 0], b[..., 1], b[..., 2])
    rs = f32(f32(ra) + f32(rb))
    return f32(d - f32(rs * rs)) < 0


def ref_free(prog, spheres, env_checks, self_checks, scene, q):
    """Per-pose verdict of the reference's generated checker (True = free)."""
    c = F.sphere_centres(prog, spheres, q, mode="ref")
    radii = [float(s["args"][3].value) for s in spheres]
    coll = np.zeros(len(q), dtype=bool)
    for chk in env_checks:
        g = chk["gate"]
        gh = env_hit(scene, c[:, g, 0], c[:, g, 1], c[:, g, 2], radii[g])
        if chk["direct"]:
            coll |= gh
            continue
        leaf = np.zeros(len(q), dtype=bool)
        for l in chk["leaves"]:
            leaf |= env_hit(scene, c[:, l, 0], c[:, l, 1], c[:, l, 2], radii[l])
        coll |= gh & leaf
    for chk in self_checks:
        ga, gb = chk["gate"]
        gh = self_hit(c[:, ga], c[:, gb], radii[ga], radii[gb])
        if chk["direct"]:
            coll |= gh
            continue
        pair = np.zeros(len(q), dtype=bool)
        for a, b in chk["pairs"]:
            pair |= self_hit(c[:, a], c[:, b], radii[a], radii[b])
        coll |= gh & pair
    return ~coll


def verdict_scenes(robot: str):
    zc = S.Scene.build(
        spheres=[S.sphere((0.55, 0.1, 0.5), 0.08), S.sphere((-0.3, 0.5, 0.9), 0.12)],
        z_capsules=[S.capsule_endpoints((0.5, -0.35, 0.0), (0.5, -0.35, 1.1), 0.05),
                    S.capsule_endpoints((-0.45, -0.4, 0.2), (-0.45, -0.4, 0.9), 0.07)],
        z_cuboids=[S.cuboid((0.6, 0.0, 0.4), (0.0, 0.0, 0.5), (0.15, 0.3, 0.01)),
                   S.cuboid((0.0, 0.7, 0.6), (0.0, 0.0, -0.3), (0.2, 0.1, 0.3))],
        capsules=[S.cylinder_center((0.3, 0.5, 0.7), (0.4, 0.2, 0.1), 0.04, 0.5)],
        cuboids=[S.cuboid((-0.5, 0.2, 0.3), (0.3, -0.2, 0.7), (0.1, 0.05, 0.2))],
    )
    avoid = fixed_spheres(robot)
    return {
        "spheres12": S.scene_spheres(0, avoid=avoid),
        "obb_cyl": S.scene_obb_cylinders(1, avoid=avoid),
        "shelf": S.scene_shelf(2, 24, 48, avoid=avoid),
        "ztables": zc,
    }


def main(ref_root: str):
    utils = os.path.join(ref_root, "src/planning/utils.cuh")
    robots_hh = os.path.join(ref_root, "src/planning/Robots.hh")
    for robot_i, robot in enumerate(("panda", "fetch", "baxter")):
        prog = F.parse(utils, robot)
        spheres, env_checks, self_checks, _order = F.structure(prog)
        lower, span = (np.array(v, dtype=f32) for v in read_limits(robots_hh, robot))
        rng = np.random.default_rng(1000 + robot_i)
        d = len(lower)

        q = (lower + span * rng.random((256, d), dtype=np.float32)).astype(f32)
        cen = F.sphere_centres(prog, spheres, q, mode="ref")
        radii = np.array([float(s["args"][3].value) for s in spheres], dtype=f32)
        np.savez_compressed(os.path.join(HERE, f"fk_{robot}.npz"), q=q, centres=cen, radii=radii)

        out = {}
        poses = (lower + span * rng.random((2048, d), dtype=np.float32)).astype(f32)
        out["poses"] = poses
        for name, sc in verdict_scenes(robot).items():
            for k, v in sc.to_arrays().items():
                out[f"{name}__{k}"] = v
            free = ref_free(prog, spheres, env_checks, self_checks, sc, poses)
            out[f"{name}__free"] = free
            print(f"{robot} {name}: {free.mean() * 100:.1f}% of {len(poses)} poses free")
        np.savez_compressed(os.path.join(HERE, f"verdict_{robot}.npz"), **out)

        # edges: start from free poses so that a useful share of edges is free
        sc = S.scene_obb_cylinders(0, avoid=fixed_spheres(robot))
        cand = (lower + span * rng.random((8192, d), dtype=np.float32)).astype(f32)
        cand = cand[ref_free(prog, spheres, env_checks, self_checks, sc, cand)]
        n_e = 192
        a = cand[:n_e]
        direction = rng.normal(size=(n_e, d))
        direction /= np.linalg.norm(direction, axis=1, keepdims=True)
        length = rng.uniform(0.02, 0.5, size=(n_e, 1))
        b = np.clip(a + direction * length, lower, lower + span).astype(f32)
        edges = np.stack([a, b], 1).astype(f32)
        res = {"edges": edges}
        for k, v in sc.to_arrays().items():
            res[f"scene__{k}"] = v
        for g in (32, 64, 128):
            delta = ((b - a) / f32(g)).astype(f32)
            ks = np.arange(1, g + 1, dtype=f32)
            pts = (a[:, None, :] + ks[None, :, None] * delta[:, None, :]).astype(f32)
            free = ref_free(prog, spheres, env_checks, self_checks, sc, pts.reshape(-1, d)).reshape(n_e, g)
            res[f"free_g{g}"] = free.all(1)
            print(f"{robot} edges g={g}: {res[f'free_g{g}'].mean() * 100:.1f}% free")
        np.savez_compressed(os.path.join(HERE, f"edge_{robot}.npz"), **res)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
