"""The C++ drop-in headers (include/prrtc/src/...): a harness written
against the reference's include paths and calls compiles, links against
libprrtc.so, and (on the GPU) solves; the header factory agrees with the
Python factory used for the synthetic scenes."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "include", "prrtc")
LIBDIR = os.path.join(ROOT, "prrtc_amd")


def compile_cpp(src, out, link=True):
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + INC, os.path.join(ROOT, "tests", "cpp", src),
           "-o", out]
    if link:
        cmd += ["-L" + LIBDIR, "-lprrtc", "-Wl,-rpath," + LIBDIR]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_dropin_harness_compiles_and_links(tmp_path):
    compile_cpp("drop_in.cpp", str(tmp_path / "drop_in"))


def test_header_factory_matches_python(tmp_path):
    from prrtc_amd import scenes as S
    exe = compile_cpp("factory_dump.cpp", str(tmp_path / "factory_dump"), link=False)
    lines = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    eulers = [(0.3, -0.7, 1.9), (0.0, 0.0, 0.0), (1.2, 0.4, -2.5)]
    for i, e in enumerate(eulers):
        cub = np.array(lines[2 * i].split()[1:], dtype=np.float64)
        cap = np.array(lines[2 * i + 1].split()[1:], dtype=np.float64)
        np.testing.assert_allclose(cub, S.cuboid((0.1, 0.2, 0.3), e, (0.4, 0.5, 0.6)), atol=2e-7)
        want = S.cylinder_center((0.1, 0.2, 0.3), e, 0.05, 0.4)
        np.testing.assert_allclose(cap[:7], want[:7], atol=2e-7)
        assert abs(cap[7] - want[7]) <= 1e-6 * want[7]


@pytest.mark.gpu
def test_dropin_harness_solves(tmp_path):
    exe = compile_cpp("drop_in.cpp", str(tmp_path / "drop_in"))
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    solved, cost, path_length, nodes = out.stdout.split()
    assert solved == "1" and int(nodes) == int(path_length) + 2 and float(cost) > 0
