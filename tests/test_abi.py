"""C-ABI boundary checks that need no GPU: the engine library loads, exports
every entry point include/opentm.h declares, and its host-only calls
behave as documented."""
import ctypes
import os
import re

import numpy as np
import pytest

import opentm_amd
from opentm_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"OTM_API[^;(]*?\b(otm_\w+)\s*\(", txt))
    return names


def test_header_matches_exports():
    names = _declared()
    assert len(names) >= 30
    lib = ctypes.CDLL(_capi.LIB_PATH)
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == set(_capi.SIGNATURES), names ^ set(_capi.SIGNATURES)


def test_config_defaults():
    lib = _capi.load()
    c = _capi.OtmConfig()
    lib.otm_config_init(ctypes.byref(c))
    assert list(c.reso) == [128, 128, 128]          # cmdline.cpp:6-8
    assert list(c.heat_ratio) == [1.0, 0.0]          # cmdline.cpp:12-13
    assert c.pre_sweeps == 2 and c.post_sweeps == 2
    assert c.fem_rel_thres == pytest.approx(1e-6)
    assert c.fp64_refine == 0 and c.fp16_stencils == 0 and c.coarse_sweeps == 0
    p = _capi.OtmDensityParams()
    lib.otm_density_params_init(ctypes.byref(p))
    assert p.filter_kind == 0 and p.filter_radius == 1.5 and p.filter_periodic == 0
    assert lib.otm_version() >= 100


def test_argument_errors_without_gpu():
    lib = _capi.load()
    ctx = ctypes.c_void_p()
    assert lib.otm_create(None, ctypes.byref(ctx)) == _capi.OTM_E_ARG
    assert b"null" in lib.otm_last_error()
    assert lib.otm_heat_matrix(None, None) == _capi.OTM_E_ARG
    assert lib.otm_oc_set_bisection(None, 0, 0.0) == _capi.OTM_E_ARG
    assert lib.otm_destroy(None) == _capi.OTM_OK
    with pytest.raises(opentm_amd.OtmError):
        _capi.check(lib.otm_update(None, None, 8, 0), "otm_update")


def test_workload_generators_are_seeded():
    from opentm_amd import workloads as w
    a = w.plain_density(16, 3)
    b = w.plain_density(16, 3)
    c = w.plain_density(16, 4)
    assert a.dtype == np.float32 and a.shape == (16, 16, 16)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)
    assert a.min() >= 1e-3 and a.max() <= 1.0
    s = w.oc_start_density(16, 0)
    assert abs(float(s.mean()) - 0.3) < 0.01 and float(np.abs(s - 0.3).max()) <= 0.05 + 1e-6


def test_cxx_dropin_header_compiles_and_reports_no_device(tmp_path):
    """Code written against the reference's C++ classes builds with g++
    against include/opentm/opentm.hpp and links libopentm.so; on a machine
    without a GPU the constructors throw std::runtime_error."""
    import subprocess
    exe = tmp_path / "dropin"
    lib_dir = os.path.dirname(_capi.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cxx", "dropin_check.cpp"), "-L", lib_dir, "-lopentm",
                        "-Wl,-rpath," + lib_dir, "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    if "constructed with a GPU" in r.stdout:
        pytest.skip("a GPU is present: the GPU path is covered by the gpu-marked drop-in test")
    assert "runtime_error" in r.stdout and "device" in r.stdout


def test_openTM_module_surface():
    """The pybind module exposes the reference's surface (main.cpp:90-120)."""
    from opentm_amd import openTM
    assert {"random", "P", "G", "D", "IWP", "example"} <= set(openTM.InitWay.__members__)
    assert set(openTM.Model.__members__) == {"mma", "oc"}
    h = openTM.homo()
    for name in ("setDensity", "setConfig", "optimize"):
        assert hasattr(h, name)
    assert callable(openTM.runInstance)


def test_struct_layouts_match_header():
    """Field names and order of the ctypes mirrors equal the header's."""
    import re
    hdr = open(os.path.join(ROOT, "include", "opentm.h")).read()
    for cname, py in [("otm_config", _capi.OtmConfig), ("otm_density_params", _capi.OtmDensityParams),
                      ("otm_init_params", _capi.OtmInitParams),
                      ("otm_custom_params", _capi.OtmCustomParams), ("otm_custom_report", _capi.OtmCustomReport),
                      ("otm_request", _capi.OtmRequest), ("otm_result", _capi.OtmResult),
                      ("otm_server_counters", _capi.OtmServerCounters), ("otm_tune_report", _capi.OtmTuneReport),
                      ("otm_log_params", _capi.OtmLogParams)]:
        body = re.search(r"typedef struct\s*\w*\s*\{([^}]*)\}\s*" + cname + r"\s*;", hdr).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        names = [re.split(r"[\s\[;]", d.strip().split()[-1])[0] for d in body.split(";") if d.strip()]
        assert names == [f[0] for f in py._fields_], cname


def test_density_file_round_trip(tmp_path):
    """Grid_H::writeDensity binary layout: int32 reso[3] then float32, x
    fastest; host-only calls need no GPU."""
    rho = np.random.default_rng(1).uniform(0, 1, (4, 5, 6)).astype(np.float32)  # (nz, ny, nx)
    path = tmp_path / "rho.bin"
    opentm_amd.write_density(path, rho)
    raw = path.read_bytes()
    assert np.frombuffer(raw[:12], np.int32).tolist() == [6, 5, 4]
    np.testing.assert_array_equal(np.frombuffer(raw[12:], np.float32), rho.reshape(-1))
    np.testing.assert_array_equal(opentm_amd.read_density(path), rho)
    bad = tmp_path / "short.bin"
    bad.write_bytes(raw[:40])
    with pytest.raises(opentm_amd.OtmError):
        opentm_amd.read_density(bad)


def test_governor_and_checker_sequences(tmp_path):
    """The drop-in ConvergeChecker / VolumeGovernor (Framework.cuh:312-411)
    against decisions computed by hand from the reference's rules."""
    import subprocess
    exe = tmp_path / "gov"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cxx", "governor_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60).stdout.splitlines()
    rows = {}
    for line in out:
        tag, *vals = line.split()
        rows.setdefault(tag, []).append(vals)
    # ConvergeChecker: "not converged" through iteration 11; then three
    # consecutive relative changes <= 1e-3
    assert rows["A"][0] == ["0"] * 12 + ["1", "1"]
    assert rows["B"][0] == ["0"] * 15 + ["1"]
    assert rows["D"][0] == ["0"] * 12 + ["1"]
    # VolumeGovernor, scenario 1: value < 0 cuts B = 1 - 1 * (1 - 0.2) = 0.2
    # and F = 0.8 (decrease F * D = 0.64); then value 0.3 with the bound
    # reached stalls (change < 1e-4 from the second step on, and the first
    # change -0.8 < 1e-4 counts too): the fifth stall widens B by
    # 0.3 * D * F = 0.192
    g1 = [(int(a), float(b), float(c)) for a, b, c in rows["G1"]]
    assert g1[0][0] == 0 and g1[0][1] == pytest.approx(0.2, abs=1e-6) and g1[0][2] == pytest.approx(0.64, abs=1e-6)
    for r_, b, _ in g1[1:5]:
        assert r_ == 0 and b == pytest.approx(0.2, abs=1e-6)
    assert g1[5][1] == pytest.approx(0.392, abs=1e-6)
    # scenario 2: the change alternates sign from the second check on; the
    # ninth check (eighth sign change) is not near the bound, so it cuts
    # B = 1 - (1 - 0.1) = 0.1; the tenth cuts again with D = 0 (F = 0.64)
    g2 = [(int(a), float(b), float(c)) for a, b, c in rows["G2"]]
    for r_, b, _ in g2[:8]:
        assert r_ == 0 and b == 1.0
    assert g2[8][1] == pytest.approx(0.1, abs=1e-6) and g2[8][2] == pytest.approx(0.8 * 0.9, abs=1e-6)
    assert g2[9][1] == pytest.approx(0.1, abs=1e-6) and g2[9][2] == pytest.approx(0.0, abs=1e-7)
    # scenario 3: F = 0.8^k after k cuts while above 0.05; 0.8^14 < 0.05
    assert [int(v[0]) for v in rows["G3"]] == [0] * 13 + [1]
    # scenario 4: itn == 100 cuts B = 1 - (1 - 0.25) = 0.25
    assert float(rows["G4"][0][0]) == pytest.approx(0.25, abs=1e-6)
