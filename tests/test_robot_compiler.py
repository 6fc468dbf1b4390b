"""The robot-model compiler (prrtc_amd.robots.urdf_compile): URDF subset +
sphere file -> model table.

* Panda regenerated from a URDF of its published kinematics and its sphere
  file: sphere centres within 1e-5 m of the reference-derived golden FK, and
  planted URDF mistakes (a sign, an offset, an axis) fail that bound;
* closed forms: a planar two-link arm, a prismatic joint on a tilted axis,
  a fixed joint's roll-pitch-yaw;
* the compiled frame chain against URDF's own composition rule at random
  poses, and the compiled table through the oracle's table-driven checker
  against the golden verdicts."""
import math
import os

import numpy as np
import pytest

from conftest import NEAR, ROOT, golden, golden_scene
from prrtc_amd.robots import urdf_compile as U

RDIR = os.path.join(ROOT, "prrtc_amd", "robots")


def panda_inputs():
    import json
    with open(os.path.join(RDIR, "panda.urdf")) as f:
        urdf = f.read()
    with open(os.path.join(RDIR, "panda_spheres.json")) as f:
        spec = json.load(f)
    return urdf, spec


def test_panda_regenerated_matches_golden_fk():
    urdf, spec = panda_inputs()
    m = U.compile_model(urdf, spec)
    g = golden("fk_panda.npz")
    err = float(np.abs(U.model_fk(m, g["q"]) - g["centres"]).max())
    assert err < 1e-5, err
    # the limits are the reference's (Robots.hh:17-41), read from the URDF
    from prrtc_amd.robots import load_model
    ref = load_model("panda")
    no_override = {k: v for k, v in spec.items() if k != "limits"}
    np.testing.assert_allclose(U.compile_model(urdf, no_override)["lower"], ref["lower"], atol=1e-12)
    assert m["env_checks"] == ref["env_checks"] and m["self_checks"] == ref["self_checks"]
    assert m["order"] == ref["order"] and len(m["spheres"]) == len(ref["spheres"])


@pytest.mark.parametrize("mistake", [
    ('<origin xyz="-0.0825 0.384 0" rpy="-1.5707963267948966 0 0"/>',
     '<origin xyz="-0.0825 0.384 0" rpy="1.5707963267948966 0 0"/>'),          # a sign
    ('<origin xyz="0.0825 0 0"', '<origin xyz="0 0 0"'),                          # a dropped offset
    ('<origin xyz="0.088 0 0" rpy="1.5707963267948966 0 0"/>\n    <axis xyz="0 0 1"/>',
     '<origin xyz="0.088 0 0" rpy="1.5707963267948966 0 0"/>\n    <axis xyz="0 0 -1"/>'),   # an axis
    ('<origin xyz="0 -0.316 0"', '<origin xyz="0 0.316 0"'),                       # a transposed offset
])
def test_planted_urdf_mistakes_fail(mistake):
    urdf, spec = panda_inputs()
    assert mistake[0] in urdf
    wrong = urdf.replace(mistake[0], mistake[1], 1)
    m = U.compile_model(wrong, spec)
    g = golden("fk_panda.npz")
    assert float(np.abs(U.model_fk(m, g["q"]) - g["centres"]).max()) > 1e-3


PLANAR = """<robot name="planar">
  <link name="base"/><link name="l1"/><link name="l2"/><link name="tool"/>
  <joint name="j1" type="revolute"><parent link="base"/><child link="l1"/>
    <origin xyz="0 0 0.1"/><axis xyz="0 0 1"/><limit lower="-3" upper="3"/></joint>
  <joint name="j2" type="continuous"><parent link="l1"/><child link="l2"/>
    <origin xyz="0.5 0 0"/><axis xyz="0 0 1"/></joint>
  <joint name="t" type="fixed"><parent link="l2"/><child link="tool"/><origin xyz="0.3 0 0"/></joint>
</robot>"""


def test_closed_form_planar_arm():
    spec = {"spheres": [{"link": "tool", "xyz": [0, 0, 0], "r": 0.05}, {"link": "l2", "xyz": [0.1, 0.2, 0], "r": 0.05},
                        {"link": "base", "xyz": [0, 0, 0.05], "r": 0.1}],
            "env": [{"label": "tool", "sphere": 0}]}
    m = U.compile_model(PLANAR, spec)
    assert m["dof"] == 2 and m["lower"] == [-3.0, -math.pi] and m["spheres"][2]["frame"] == -1
    rng = np.random.default_rng(0)
    q = rng.uniform(-3, 3, (50, 2))
    c = U.model_fk(m, q)
    a, b = q[:, 0], q[:, 0] + q[:, 1]
    np.testing.assert_allclose(c[:, 0, 0], 0.5 * np.cos(a) + 0.3 * np.cos(b), atol=1e-12)
    np.testing.assert_allclose(c[:, 0, 1], 0.5 * np.sin(a) + 0.3 * np.sin(b), atol=1e-12)
    np.testing.assert_allclose(c[:, 0, 2], 0.1, atol=1e-12)
    # a point off the link's x axis: (0.1, 0.2) rotated by q1 + q2 about the elbow
    np.testing.assert_allclose(c[:, 1, 0], 0.5 * np.cos(a) + 0.1 * np.cos(b) - 0.2 * np.sin(b), atol=1e-12)
    np.testing.assert_allclose(c[:, 1, 1], 0.5 * np.sin(a) + 0.1 * np.sin(b) + 0.2 * np.cos(b), atol=1e-12)
    np.testing.assert_allclose(c[:, 2], np.broadcast_to([0, 0, 0.05], (50, 3)), atol=1e-12)


def test_closed_form_prismatic_and_rpy():
    urdf = """<robot name="slide">
      <link name="base"/><link name="carriage"/><link name="tip"/>
      <joint name="s" type="prismatic"><parent link="base"/><child link="carriage"/>
        <origin xyz="1 2 3"/><axis xyz="1 1 0"/><limit lower="0" upper="0.5"/></joint>
      <joint name="f" type="fixed"><parent link="carriage"/><child link="tip"/>
        <origin xyz="0 0 0" rpy="0.3 -0.7 1.9"/></joint>
    </robot>"""
    m = U.compile_model(urdf, {"spheres": [{"link": "tip", "xyz": [1, 0, 0], "r": 0.1}]})
    assert m["joints"][0]["type"] == "prismatic"
    q = np.array([[0.0], [0.2], [0.5]])
    c = U.model_fk(m, q)[:, 0]
    # tip x axis: the first column of Rz(1.9) Ry(-0.7) Rx(0.3)
    cr, cp, cy, sy, sp = math.cos(0.3), math.cos(-0.7), math.cos(1.9), math.sin(1.9), math.sin(-0.7)
    col = np.array([cy * cp, sy * cp, -sp])
    want = np.array([1, 2, 3]) + q * np.array([1, 1, 0]) / math.sqrt(2) + col
    np.testing.assert_allclose(c, want, atol=1e-12)


def test_compiled_chain_equals_urdf_composition():
    """Frames aligned to the joint axes and fixed joints folded in give the
    same world spheres as URDF's own link-by-link composition."""
    urdf, spec = panda_inputs()
    m = U.compile_model(urdf, spec)
    rng = np.random.default_rng(3)
    lo, rg = np.array(m["lower"]), np.array(m["range"])
    for q in lo + rg * rng.random((8, 7)):
        poses = U.urdf_link_poses(urdf, q)
        want = np.array([poses[s["link"]][:3, :3] @ s["xyz"] + poses[s["link"]][:3, 3] for s in spec["spheres"]])
        np.testing.assert_allclose(U.model_fk(m, q)[0], want, atol=1e-12)


def test_compiled_table_verdicts(oracle_robots):
    """The compiled table drives the oracle's table-driven checker (the same
    tables the device code is generated from) to the reference's verdicts."""
    from oracle import oracle
    urdf, spec = panda_inputs()
    R = oracle.OracleRobot(U.compile_model(urdf, spec))
    g = golden("verdict_panda.npz")
    for scene in ("spheres12", "obb_cyl", "shelf"):
        sc = golden_scene(g, scene)
        free, margin = R.check(sc, g["poses"], margin=True)
        assert not np.any((free != g[f"{scene}__free"]) & (margin >= NEAR)), scene


def test_gen_headers_emits_compiled_table():
    from prrtc_amd.robots import gen_headers
    urdf, spec = panda_inputs()
    code = gen_headers.gen_robot_from(U.compile_model(urdf, spec))
    assert "PRRTC_FK_BODY" in code and code.count("sincosf") == 7


def test_rejects_outside_subset():
    with pytest.raises(ValueError, match="subset"):
        U.parse_urdf('<robot name="x"><link name="a"/><link name="b"/><joint name="j" type="floating">'
                     '<parent link="a"/><child link="b"/></joint></robot>')
    with pytest.raises(ValueError, match="no limits"):
        U.compile_model('<robot name="x"><link name="a"/><link name="b"/><joint name="j" type="revolute">'
                        '<parent link="a"/><child link="b"/></joint></robot>', {"spheres": []})


# ---------------------------------------------------------------- UR5
# A fourth robot, not in the reference, compiled from a URDF of the UR5's
# published kinematics and this project's sphere file.

def ur5_inputs():
    import json
    with open(os.path.join(RDIR, "ur5.urdf")) as f:
        urdf = f.read()
    with open(os.path.join(RDIR, "ur5_spheres.json")) as f:
        spec = json.load(f)
    return urdf, spec


def test_ur5_closed_form_tool_point():
    """At the zero pose the tool frame sits at (0.81725, 0.19145, -0.005491)
    (0.425 + 0.39225 out, 0.13585 - 0.1197 + 0.10915 + 0.0823 across,
    0.089159 - 0.09465 up); the ee sphere is 0.04 m along the tool's x axis,
    which points along +y there.  Turning the base joint by q turns it about z."""
    urdf, spec = ur5_inputs()
    m = U.compile_model(urdf, spec)
    ee = next(i for i, s in enumerate(spec["spheres"]) if s["link"] == "ee_link")
    tool = np.array([0.81725, 0.19145 + 0.04, 0.089159 - 0.09465])
    np.testing.assert_allclose(U.model_fk(m, np.zeros((1, 6)))[0, ee], tool, atol=1e-9)
    for q0 in (0.7, -2.1):
        c, s = math.cos(q0), math.sin(q0)
        want = np.array([c * tool[0] - s * tool[1], s * tool[0] + c * tool[1], tool[2]])
        np.testing.assert_allclose(U.model_fk(m, np.array([[q0, 0, 0, 0, 0, 0]]))[0, ee], want, atol=1e-9)


def test_ur5_compiled_chain_equals_urdf_composition():
    urdf, spec = ur5_inputs()
    m = U.compile_model(urdf, spec)
    rng = np.random.default_rng(5)
    lo, rg = np.array(m["lower"]), np.array(m["range"])
    for q in lo + rg * rng.random((8, 6)):
        poses = U.urdf_link_poses(urdf, q)
        want = np.array([poses[s["link"]][:3, :3] @ s["xyz"] + poses[s["link"]][:3, 3] for s in spec["spheres"]])
        np.testing.assert_allclose(U.model_fk(m, q)[0][:len(want)], want, atol=1e-12)


def test_ur5_committed_table_and_header_are_the_compilers_output():
    """robots/ur5.json is compile_model's output and csrc/robots/ur5.gen.h is
    gen_headers' output for it (the device code of the fourth robot)."""
    import json
    from prrtc_amd.robots import gen_headers, load_model
    urdf, spec = ur5_inputs()
    m = U.compile_model(urdf, spec)
    assert json.loads(json.dumps(m)) == load_model("ur5")
    with open(os.path.join(ROOT, "prrtc_amd", "csrc", "robots", "ur5.gen.h")) as f:
        assert f.read() == gen_headers.gen_robot("ur5")


def test_ur5_oracle_verdicts_brute_force(oracle_robots):
    """The oracle's table-driven checker on the compiled UR5 against a brute
    force over every sphere of its env groups and self pairs (sphere
    obstacles, float64 FK), away from contact."""
    from oracle import oracle
    from prrtc_amd import scenes as S
    from prrtc_amd.robots import load_model
    m = load_model("ur5")
    R = oracle.OracleRobot(m)
    rng = np.random.default_rng(11)
    obst = np.c_[rng.uniform(-0.9, 0.9, (10, 2)), rng.uniform(0.0, 0.9, 10), rng.uniform(0.05, 0.15, 10)]
    sc = S.Scene.build(spheres=[S.sphere(o[:3], o[3]) for o in obst])
    lo, rg = np.array(m["lower"]), np.array(m["range"])
    q = (lo + rg * rng.random((400, 6))).astype(np.float32)
    free, margin = R.check(sc, q, margin=True)
    c = U.model_fk(m, q.astype(np.float64))
    rad = np.array([s["r"] for s in m["spheres"]])
    env_sph = sorted({i for g in m["env_checks"] for i in (g["leaves"] or [g["gate"]])})
    pairs = [g["gate"] if g["direct"] else None for g in m["self_checks"]]
    assert all(p is not None for p in pairs)
    d = np.linalg.norm(c[:, env_sph, None, :] - obst[None, None, :, :3], axis=-1) - rad[env_sph][None, :, None] - obst[None, None, :, 3]
    hit_env = (d < 0).any(axis=(1, 2))
    dp = np.stack([np.linalg.norm(c[:, a] - c[:, b], axis=-1) - rad[a] - rad[b] for a, b in pairs], axis=1)
    hit_self = (dp < 0).any(axis=1)
    brute_free = ~(hit_env | hit_self)
    far = np.minimum(np.abs(d).min(axis=(1, 2)), np.abs(dp).min(axis=1)) > 1e-5
    assert far.sum() > 300 and 0.1 < brute_free.mean() < 0.9, (far.sum(), brute_free.mean())
    assert np.array_equal(free[far].astype(bool), brute_free[far])
