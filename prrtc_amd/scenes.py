"""Obstacle construction and seeded synthetic scenes.

Obstacle constructors follow the reference's factory (src/collision/factory.hh):

* cuboid from centre, Euler XYZ angles and half-extents, with
  ``R = Rz(phi) * Ry(theta) * Rx(rho)`` and the box axes the columns of R
  (factory.hh:28-63);
* cylinder from centre, Euler XYZ angles, radius and length: endpoints at
  ``centre +- R * (0, 0, L/2)`` (factory.hh:165-185), stored as a capsule
  ``(p1, v = p2 - p1, r, rdv = 1/|v|^2)`` (factory.hh:118-129);
* sphere ``(x, y, z, r)`` (factory.hh:351-355).

The reference's checker tests spheres, capsules, z-aligned capsules, cuboids
and z-aligned cuboids (src/planning/utils.cuh:211-242); the reference's MBM
loader turns every cylinder into a capsule (scripts/evaluate_mbm.cpp:41-52).
A ``Scene`` therefore holds those five tables, as float32 arrays:

    spheres    [n, 4]   x y z r
    capsules   [n, 8]   x1 y1 z1 xv yv zv r rdv
    z_capsules [n, 8]   same layout, v along z
    cuboids    [n, 15]  x y z  a1x a1y a1z  a2x a2y a2z  a3x a3y a3z  r1 r2 r3
    z_cuboids  [n, 15]  same layout, a3 = z

Synthetic scenes use SplitMix64 (Steele, Lea & Flood 2014) for all random
numbers, Shoemake's method (1992) for uniform unit quaternions, and
exponential inter-arrival times for Poisson streams.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32
MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator; ``uniform()`` returns 24-bit floats in [0, 1)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 40) * (1.0 / 16777216.0)

    def range(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.uniform()

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + (self.next_u64() % (hi - lo + 1))

    def quaternion(self):
        """Uniform random unit quaternion (w, x, y, z), Shoemake 1992."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
        t1, t2 = 2.0 * math.pi * u2, 2.0 * math.pi * u3
        return (b * math.cos(t2), a * math.sin(t1), a * math.cos(t1), b * math.sin(t2))


def euler_xyz_matrix(rho: float, theta: float, phi: float) -> np.ndarray:
    """R = Rz(phi) Ry(theta) Rx(rho) in float32 (factory.hh:39-45)."""
    cr, sr = math.cos(rho), math.sin(rho)
    ct, st = math.cos(theta), math.sin(theta)
    cp, sp = math.cos(phi), math.sin(phi)
    R = np.array([
        [cp * ct, cp * st * sr - sp * cr, cp * st * cr + sp * sr],
        [sp * ct, sp * st * sr + cp * cr, sp * st * cr - cp * sr],
        [-st, ct * sr, ct * cr],
    ])
    return R.astype(F32)


def quat_matrix(w, x, y, z) -> np.ndarray:
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ]).astype(F32)


def cuboid_from_matrix(center, R, half) -> np.ndarray:
    R = np.asarray(R, dtype=F32)
    return np.array([center[0], center[1], center[2],
                     R[0, 0], R[1, 0], R[2, 0],
                     R[0, 1], R[1, 1], R[2, 1],
                     R[0, 2], R[1, 2], R[2, 2],
                     half[0], half[1], half[2]], dtype=F32)


def cuboid(center, euler_xyz, half) -> np.ndarray:
    """factory::cuboid::array (factory.hh:28-63, 95-109)."""
    return cuboid_from_matrix(center, euler_xyz_matrix(*euler_xyz), half)


def capsule_endpoints(p1, p2, radius) -> np.ndarray:
    """factory::cylinder::endpoints::flat (factory.hh:118-129)."""
    p1 = np.asarray(p1, dtype=F32)
    p2 = np.asarray(p2, dtype=F32)
    v = (p2 - p1).astype(F32)
    dot = F32(F32(v[0] * v[0]) + F32(v[1] * v[1]) + F32(v[2] * v[2]))
    rdv = F32(1.0 / float(dot))
    return np.array([p1[0], p1[1], p1[2], v[0], v[1], v[2], radius, rdv], dtype=F32)


def cylinder_center(center, euler_xyz, radius, length) -> np.ndarray:
    """factory::cylinder::center::flat (factory.hh:165-185) -> capsule row."""
    R = euler_xyz_matrix(*euler_xyz)
    half = F32(length / 2.0)
    c = np.asarray(center, dtype=F32)
    p1 = c + R[:, 2] * half
    p2 = c - R[:, 2] * half
    return capsule_endpoints(p1, p2, radius)


def sphere(center, radius) -> np.ndarray:
    return np.array([center[0], center[1], center[2], radius], dtype=F32)


@dataclass
class Scene:
    spheres: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), F32))
    capsules: np.ndarray = field(default_factory=lambda: np.zeros((0, 8), F32))
    z_capsules: np.ndarray = field(default_factory=lambda: np.zeros((0, 8), F32))
    cuboids: np.ndarray = field(default_factory=lambda: np.zeros((0, 15), F32))
    z_cuboids: np.ndarray = field(default_factory=lambda: np.zeros((0, 15), F32))

    @staticmethod
    def build(spheres=(), capsules=(), z_capsules=(), cuboids=(), z_cuboids=()):
        def arr(rows, w):
            rows = list(rows)
            return np.ascontiguousarray(np.array(rows, dtype=F32).reshape(len(rows), w))
        return Scene(arr(spheres, 4), arr(capsules, 8), arr(z_capsules, 8), arr(cuboids, 15), arr(z_cuboids, 15))

    def counts(self):
        return (len(self.spheres), len(self.capsules), len(self.z_capsules), len(self.cuboids), len(self.z_cuboids))

    def n_obstacles(self) -> int:
        return sum(self.counts())

    def to_arrays(self) -> dict:
        return {"spheres": self.spheres, "capsules": self.capsules, "z_capsules": self.z_capsules,
                "cuboids": self.cuboids, "z_cuboids": self.z_cuboids}

    @staticmethod
    def from_arrays(d) -> "Scene":
        return Scene(*(np.ascontiguousarray(np.asarray(d[k], dtype=F32)) for k in
                       ("spheres", "capsules", "z_capsules", "cuboids", "z_cuboids")))


# --------------------------------------------------------------------------
# synthetic workloads

def obstacle_distance(kind: str, row, p) -> float:
    """Euclidean distance from point p to an obstacle row (0 inside)."""
    p = np.asarray(p, dtype=np.float64)
    row = np.asarray(row, dtype=np.float64)
    if kind == "sphere":
        return max(0.0, float(np.linalg.norm(p - row[:3]) - row[3]))
    if kind in ("capsule", "z_capsule"):
        p1, v, r = row[:3], row[3:6], row[6]
        t = np.clip(np.dot(p - p1, v) / max(np.dot(v, v), 1e-30), 0.0, 1.0)
        return max(0.0, float(np.linalg.norm(p - (p1 + t * v)) - r))
    axes = row[3:12].reshape(3, 3)
    d = p - row[:3]
    ex = np.maximum(np.abs(axes @ d) - row[12:15], 0.0)
    return float(np.linalg.norm(ex))


def _clear_of(kind: str, row, avoid, margin: float = 0.02) -> bool:
    for s in avoid:
        if obstacle_distance(kind, row, s[:3]) < s[3] + margin:
            return False
    return True


def _centre(rng: SplitMix64, keep_out: float = 0.25):
    """Centre uniform in [-0.8,0.8]^2 x [0,1.2], at least keep_out from the z axis."""
    while True:
        x, y, z = rng.range(-0.8, 0.8), rng.range(-0.8, 0.8), rng.range(0.0, 1.2)
        if math.hypot(x, y) >= keep_out:
            return (x, y, z)


def _draw(rng, make, kind, avoid):
    """Draw an obstacle with ``make(rng)``; redraw while it touches ``avoid``
    (the robot's fixed base spheres, which no motion can clear)."""
    for _ in range(10000):
        row = make(rng)
        if avoid is None or len(avoid) == 0 or _clear_of(kind, row, avoid):
            return row
    raise RuntimeError("could not place an obstacle clear of the robot base")


def _rand_sphere(rng):
    return sphere(_centre(rng), rng.range(0.05, 0.15))


def _rand_obb(rng):
    c = _centre(rng)
    half = (rng.range(0.02, 0.2), rng.range(0.02, 0.2), rng.range(0.02, 0.2))
    return cuboid_from_matrix(c, quat_matrix(*rng.quaternion()), half)


def _rand_vertical_cylinder(rng):
    c = _centre(rng)
    return cylinder_center(c, (0.0, 0.0, 0.0), rng.range(0.03, 0.08), rng.range(0.2, 0.6))


def scene_spheres(seed: int, n: int = 12, avoid=None) -> Scene:
    """12 spheres, r ~ U(0.05, 0.15), centres in the box, 0.25 m off the base axis."""
    rng = SplitMix64(seed)
    return Scene.build(spheres=[_draw(rng, _rand_sphere, "sphere", avoid) for _ in range(n)])


def scene_obb_cylinders(seed: int, n_obb: int = 24, n_cyl: int = 8, avoid=None) -> Scene:
    """24 oriented cuboids (half-extents U(0.02,0.2), uniform orientation) and
    8 vertical cylinders (r ~ U(0.03,0.08), h ~ U(0.2,0.6)) as capsules."""
    rng = SplitMix64(seed)
    obbs = [_draw(rng, _rand_obb, "cuboid", avoid) for _ in range(n_obb)]
    cyls = [_draw(rng, _rand_vertical_cylinder, "capsule", avoid) for _ in range(n_cyl)]
    return Scene.build(cuboids=obbs, capsules=cyls)


def _mixed(rng, n, avoid):
    sph, obb, cyl = [], [], []
    for _ in range(n):
        u = rng.uniform()
        if u < 0.5:
            sph.append(_draw(rng, _rand_sphere, "sphere", avoid))
        elif u < 0.8:
            obb.append(_draw(rng, _rand_obb, "cuboid", avoid))
        else:
            cyl.append(_draw(rng, _rand_vertical_cylinder, "capsule", avoid))
    return Scene.build(spheres=sph, cuboids=obb, capsules=cyl)


def scene_mixed(seed: int, n_min: int = 8, n_max: int = 64, avoid=None) -> Scene:
    """Obstacle count U{n_min..n_max}: 50% spheres, 30% OBBs, 20% cylinders."""
    rng = SplitMix64(seed)
    return _mixed(rng, rng.integer(n_min, n_max), avoid)


def scene_shelf(seed: int, n_min: int = 4, n_max: int = 128, avoid=None) -> Scene:
    """Shelf: 3-5 axis-aligned boards 0.02 m thick spaced U(0.25,0.4) apart
    vertically, plus random spheres; total count log-uniform in [n_min, n_max]."""
    rng = SplitMix64(seed)
    n_total = int(round(math.exp(rng.range(math.log(n_min), math.log(n_max)))))
    n_boards = rng.integer(3, 5)
    # shelf placed beside the robot, its boards axis-aligned about z
    while True:
        ang = rng.range(-math.pi, math.pi)
        dist = rng.range(0.5, 0.8)
        cx, cy = dist * math.cos(ang), dist * math.sin(ang)
        width, depth = rng.range(0.3, 0.6), rng.range(0.2, 0.35)
        z = rng.range(0.05, 0.3)
        R = euler_xyz_matrix(0.0, 0.0, ang)
        boards = []
        for _ in range(n_boards):
            boards.append(cuboid_from_matrix((cx, cy, z), R, (depth / 2, width / 2, 0.01)))
            z += rng.range(0.25, 0.4)
        if avoid is None or all(_clear_of("cuboid", b, avoid) for b in boards):
            break
    spheres_ = [_draw(rng, _rand_sphere, "sphere", avoid) for _ in range(max(0, n_total - n_boards))]
    return Scene.build(spheres=spheres_, z_cuboids=boards)


def scene_serving(seed: int, avoid=None) -> Scene:
    """4-128 mixed primitives (log-uniform count): spheres, OBBs, cylinders."""
    rng = SplitMix64(seed)
    n = int(round(math.exp(rng.range(math.log(4), math.log(128)))))
    return _mixed(rng, n, avoid)


def uniform_poses(rng: SplitMix64, lower, span, n: int) -> np.ndarray:
    lower = np.asarray(lower, dtype=F32)
    span = np.asarray(span, dtype=F32)
    u = np.array([[rng.uniform() for _ in range(len(lower))] for _ in range(n)], dtype=F32)
    return (u * span + lower).astype(F32)


def random_edges(rng: SplitMix64, lower, span, n: int, max_len: float = 0.5) -> np.ndarray:
    """Edges (a, b): a uniform in the limits, b = a + L u, u uniform on the unit
    sphere, L ~ U(0, max_len], resampled while b leaves the limits."""
    lower = np.asarray(lower, dtype=np.float64)
    upper = lower + np.asarray(span, dtype=np.float64)
    d = len(lower)
    out = np.zeros((n, 2, d), dtype=F32)
    for i in range(n):
        a = lower + (upper - lower) * np.array([rng.uniform() for _ in range(d)])
        while True:
            g = np.array([_gauss(rng) for _ in range(d)])
            g /= np.linalg.norm(g)
            L = max_len * (1.0 - rng.uniform())
            b = a + L * g
            if np.all(b >= lower) and np.all(b <= upper):
                break
        out[i, 0] = a
        out[i, 1] = b
    return out


def _gauss(rng: SplitMix64) -> float:
    u1 = 1.0 - rng.uniform()
    u2 = rng.uniform()
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
