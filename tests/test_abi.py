"""The C-ABI library loads and exports every symbol include/prrtc.h declares;
host-side logic that needs no GPU (settings defaults, scene packing)."""
import ctypes as C
import os
import re

import numpy as np

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "prrtc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prrtc_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import prrtc_amd.lib as L
    lib = C.CDLL(L.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 12
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes binding covers the same set
    assert sorted(name for name, _, _ in L.SIGNATURES) == syms


def test_settings_default_match_reference():
    import prrtc_amd
    d = prrtc_amd.settings_default()
    # pRRTC_settings.hh:4-16
    assert d["max_samples"] == 1000000 and d["max_iters"] == 1000000
    assert d["num_new_configs"] == 600 and d["granularity"] == 64
    assert d["range"] == np.float32(0.5) and d["balance"] == 1 and d["tree_ratio"] == np.float32(0.5)
    assert d["dynamic_domain"] == 1 and d["dd_alpha"] == np.float32(1e-4)
    assert d["dd_radius"] == 4.0 and d["dd_min_radius"] == 1.0 and d["seed"] == 1


def test_robot_tables():
    import prrtc_amd
    L = prrtc_amd.load()
    assert [L.prrtc_robot_dof(i) for i in range(4)] == [3, 7, 8, 14]
    assert [L.prrtc_robot_num_spheres(i) for i in range(4)] == [1, 69, 125, 100]
    assert L.prrtc_robot_dof(7) == -1


def test_scene_pack_layout():
    import prrtc_amd
    from prrtc_amd import scenes as S
    sc = S.Scene.build(spheres=[S.sphere((1, 2, 3), 0.5)],
                       capsules=[S.capsule_endpoints((0, 0, 0), (0, 0, 2), 0.1)],
                       cuboids=[S.cuboid((1, 1, 1), (0, 0, 0), (0.1, 0.2, 0.3))])
    p = prrtc_amd.pack_scene(sc)
    hdr = p[:8].view(np.int32)
    assert list(hdr[:6]) == [1, 1, 0, 1, 0, 1 + 2 + 4]
    assert len(p) == 8 + 4 * 7
    np.testing.assert_array_equal(p[8:12], [1, 2, 3, 0.5])
    np.testing.assert_allclose(p[12:20], [0, 0, 0, 0.1, 0, 0, 2, 0.25])
    # the last float of a cuboid is its cull radius: half-diagonal padded by 2^-16 and 1 mm
    cull = np.sqrt(0.1 ** 2 + 0.2 ** 2 + 0.3 ** 2) * (1 + 2.0 ** -16) + 1e-3
    np.testing.assert_allclose(p[20:36], [1, 1, 1, 0.1, 1, 0, 0, 0.2, 0, 1, 0, 0.3, 0, 0, 1, cull], atol=1e-7)


def test_invalid_arguments_fail_without_gpu():
    """Argument checks run before any device work and return EINVAL."""
    import prrtc_amd
    from prrtc_amd.lib import CScene, load
    L = load()
    bad = CScene()
    bad.n_spheres = 3          # count without data
    rc = L.prrtc_check_poses(1, C.byref(bad), None, 0, None)
    assert rc == -1
    assert b"invalid" in L.prrtc_last_error()
    assert L.prrtc_validate_edges(9, C.byref(CScene()), None, 0, 32, None) == -1
    # nearest-node search: cap must be whole 1024-node chunks, dof one of the robots'
    assert L.prrtc_nearest_async(7, 1, 1000, 10, 1, 1, 1, 1, 1, 1, None) == -1
    assert L.prrtc_nearest_async(5, 1, 1024, 10, 1, 1, 1, 1, 1, 1, None) == -1
    assert L.prrtc_nearest_async(7, 1, 1024, 2000, 1, 1, 1, 1, 1, 1, None) == -1
    assert L.prrtc_nearest_async(7, 1, 1024, 10, 1, 1, 1, 9, 1, 1, None) == -1
    # batch entry points: negative count, settings count neither 1 nor n
    for fn in (L.prrtc_solve_batch, L.prrtc_solve_batch_rounds):
        assert fn(-1, None, None, 1, None, None, 16, None) == -1
        assert fn(2, None, None, 3, None, None, 16, None) == -1
    # tuning ranges
    from prrtc_amd.lib import CTuning
    assert L.prrtc_set_tuning(1, C.byref(CTuning(0, 9, 0))) == -1
    assert L.prrtc_set_tuning(7, C.byref(CTuning(0, 0, 0))) == -1
    t = CTuning()
    assert L.prrtc_get_tuning(2, C.byref(t)) == 0 and t.crowd_above == 0   # 0: the default threshold


def test_server_rejects_bad_config_without_gpu():
    from prrtc_amd.lib import CServerConfig, load
    L = load()
    h = C.c_void_p()
    cfg = CServerConfig(0, -5, 0, 0, 0, 0)   # negative table size
    assert L.prrtc_server_create(C.byref(cfg), C.byref(h)) == -1
    assert not h.value
    assert L.prrtc_server_create(None, None) == -1
    # path pool of max_resident x (path_cap - 2) rows: 2^20 x 4094 > INT32_MAX
    cfg = CServerConfig(0, 1 << 20, 0, 0, 4096, 0)
    assert L.prrtc_server_create(C.byref(cfg), C.byref(h)) == -1 and not h.value
    assert b"path rows" in L.prrtc_last_error()


def test_cuboid_cull_radius():
    """The pack's cull radius bounds the box: at least its half-diagonal,
    at most that plus the stated pads (2^-16 relative, 1 mm)."""
    import prrtc_amd
    from prrtc_amd import scenes as S
    rng = np.random.default_rng(3)
    for _ in range(50):
        h = rng.uniform(0.001, 2.0, 3).astype(np.float32)
        sc = S.Scene.build(cuboids=[S.cuboid(rng.uniform(-1, 1, 3), rng.uniform(-3, 3, 3), h)])
        cull = float(prrtc_amd.pack_scene(sc)[8 + 15])
        R = float(np.linalg.norm(h.astype(np.float64)))
        assert R + 1e-3 <= cull * (1 + 1e-7)
        assert cull <= R * (1 + 2.0 ** -16) + 1e-3 + 1e-6


def test_ctypes_structs_match_the_header(tmp_path):
    """Every ctypes mirror in prrtc_amd/lib.py has the size and field offsets
    the C compiler gives the header's struct (a mismatch garbles every
    result after the first of a batch)."""
    import subprocess
    import prrtc_amd.lib as L
    exe = str(tmp_path / "abi_sizes")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "abi_sizes.c"), "-o", exe], check=True)
    got = dict(l.split() for l in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines())
    mirrors = {"prrtc_scene": L.CScene, "prrtc_settings": L.CSettings, "prrtc_result": L.CResult,
               "prrtc_problem": L.CProblem, "prrtc_batch_stats": L.CBatchStats, "prrtc_tuning": L.CTuning,
               "prrtc_server_config": L.CServerConfig, "prrtc_server_stats": L.CServerStats,
               "prrtc_request_timing": L.CRequestTiming}
    for name, cls in mirrors.items():
        assert int(got[name]) == C.sizeof(cls), name
    for key, v in got.items():
        if "." in key:
            name, field = key.split(".")
            assert getattr(mirrors[name], field).offset == int(v), key
