"""C-ABI checks that need no GPU.

* the library loads and exports every function include/cgfdm.h declares, and
  the Python wrapper binds exactly that set;
* parameter parsing of the reference's own job files (tests/golden/params_*.json,
  copied from the reference's jobs/*/params.json) follows set_params.cu:
  defaults, the "rho1" key landing in bi_rho1 (set_params.cu:247-248), the
  derived ni..nz and NT = (int)(TMAX/DT) in single precision
  (set_params.cu:286-295);
* error codes for a bad file, bad JSON and unknown keys.
"""
import ctypes
import glob
import json
import os
import re

import numpy as np
import pytest

import dynfd3dcuda_amd as dyn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "cgfdm.h")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def lib():
    try:
        return dyn.load_library()
    except OSError:
        from dynfd3dcuda_amd import build
        build.build()
        return dyn.load_library()


def header_functions():
    txt = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(cgfdm_\w+)\s*\(", txt)))


def test_library_exports_every_header_function():
    L = lib()
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(dyn.EXPORTED) == names


def create(text):
    L = lib()
    h = ctypes.c_void_p()
    rc = L.cgfdm_create_from_string(text.encode(), ctypes.byref(h))
    return rc, h


def get(h, key):
    v = ctypes.c_double()
    rc = lib().cgfdm_get_param(h, key.encode(), ctypes.byref(v))
    assert rc == 0, (key, lib().cgfdm_last_error())
    return v.value


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "params_*.json"))))
def test_reference_job_params(path):
    text = open(path).read()
    js = json.loads(text)
    rc, h = create(text)
    assert rc == 0, lib().cgfdm_last_error()
    try:
        f32 = lambda x: float(np.float32(x))
        for k in ("NX", "NY", "NZ", "PX", "PY", "PZ", "INPORT_GRID_TYPE", "INPORT_STRESS_TYPE", "Friction_type",
                  "EXPORT_TIME_SKIP", "EXPORT_WAVE_SLICE_X", "EXPORT_WAVE_SLICE_Y", "EXPORT_WAVE_SLICE_Z", "PML_N",
                  "DAMP_N"):
            if k in js:
                assert get(h, k) == int(js[k]), k
        for k in ("DH", "mu_s", "mu_d", "Dc", "C0", "PML_velocity", "PML_bmax", "PML_fc", "bi_vp1", "bi_vs1",
                  "bi_vp2", "bi_vs2", "bi_rho2", "vp1", "vs1"):
            if k in js:
                assert get(h, k) == pytest.approx(js[k], rel=1e-6), k
        # "rho1" is read into bi_rho1 before bi_rho1 itself (set_params.cu:247-254)
        want_bi_rho1 = js.get("bi_rho1", js.get("rho1"))
        if want_bi_rho1 is not None:
            assert get(h, "bi_rho1") == pytest.approx(want_bi_rho1)
        for a in range(4):
            assert get(h, f"Fault_grid{a}") == js["Fault_grid"][a]
        ni, nj, nk = js["NX"] // js["PX"], js["NY"] // js["PY"], js["NZ"] // js["PZ"]
        assert (get(h, "ni"), get(h, "nj"), get(h, "nk")) == (ni, nj, nk)
        assert (get(h, "nx"), get(h, "ny"), get(h, "nz")) == (ni + 6, nj + 6, nk + 6)
        nt = int(np.float32(np.float32(js["TMAX"]) / np.float32(js["DT"])))
        assert get(h, "NT") == nt
    finally:
        lib().cgfdm_destroy(h)


def test_defaults_match_reference():
    rc, h = create('{"NX": 10, "NY": 12, "NZ": 14, "TMAX": 1.0, "DT": 0.01, "DH": 100}')
    assert rc == 0
    try:
        # PML defaults set before reading the keys (set_params.cu:263-267)
        assert get(h, "PML_N") == 20
        assert get(h, "PML_velocity") == 3000.0
        assert get(h, "PML_bmax") == 3.0
        assert get(h, "PML_fc") == 1.0
        assert get(h, "NT") == 100
    finally:
        lib().cgfdm_destroy(h)


def test_error_codes():
    L = lib()
    h = ctypes.c_void_p()
    assert L.cgfdm_create(b"/nonexistent/params.json", ctypes.byref(h)) == dyn.EFILE
    assert L.cgfdm_create_from_string(b'{"NX": 10,', ctypes.byref(h)) == dyn.EJSON
    rc, h = create('{"NX": 10, "some_unused_key": [1, 2, {"a": "b"}], "OUT": "out"}')
    assert rc == 0  # unknown keys are ignored like cJSON lookups ignore them
    try:
        assert L.cgfdm_set_param(h, b"no_such_param", ctypes.c_double(1.0)) == dyn.EPARAM
        assert L.cgfdm_set_param(h, b"NX", ctypes.c_double(32.0)) == 0
        assert get(h, "NX") == 32
        v = ctypes.c_double()
        assert L.cgfdm_get_param(h, b"no_such_param", ctypes.byref(v)) == dyn.EPARAM
        # stepping before setup is a state error, not a crash
        assert L.cgfdm_step(h, 0) == dyn.ESTATE
    finally:
        L.cgfdm_destroy(h)


def test_python_wrapper_flags_table():
    # FD_Flags (main.cu:86-94): eight distinct sign triples, each with its negation present
    s = set(dyn.FD_FLAGS)
    assert len(s) == 8
    assert all(tuple(-x for x in t) in s for t in s)


def test_cli_missing_params_file(tmp_path):
    """The command-line driver ends with -1 when the params file is missing
    (set_params.cu:108-132); nothing touches a GPU before that."""
    import subprocess
    cli = os.path.join(ROOT, "dynfd3dcuda_amd", "cgfdm")
    if not os.path.exists(cli):
        pytest.skip("cgfdm not built")
    r = subprocess.run([cli, "no_such_params.json"], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode == 255, (r.returncode, r.stdout)
    assert "no_such_params.json" in r.stdout


def test_stage_flags_table():
    """cgfdm_stage_flags (no GPU needed): row it % 8 of the reference's table
    (main.cu:86-94), negated on odd stages; bad indices rejected."""
    table = [(-1, -1, -1), (1, 1, -1), (1, 1, 1), (-1, -1, 1),
             (-1, 1, -1), (1, -1, -1), (1, -1, 1), (-1, 1, 1)]
    L = lib()
    f = [ctypes.c_int() for _ in range(3)]
    for it in range(17):
        for irk in range(4):
            assert L.cgfdm_stage_flags(it, irk, *[ctypes.byref(x) for x in f]) == 0
            sign = -1 if irk % 2 else 1
            assert tuple(x.value for x in f) == tuple(sign * v for v in table[it % 8])
    assert L.cgfdm_stage_flags(0, 4, *[ctypes.byref(x) for x in f]) == dyn.EPARAM
    assert L.cgfdm_stage_flags(-1, 0, *[ctypes.byref(x) for x in f]) == dyn.EPARAM


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "params_*.json"))))
def test_print_params_block(path, capfd):
    """cgfdm_print_params writes the reference's parameter block
    (set_params.cu:299-388) for each of the reference's job files: the
    values it read, in the reference's line format."""
    js = json.loads(open(path).read())
    rc, h = create(open(path).read())
    assert rc == 0
    try:
        capfd.readouterr()
        assert lib().cgfdm_print_params(h) == 0
        out = capfd.readouterr().out
    finally:
        lib().cgfdm_destroy(h)
    g = lambda x: "%g" % x  # noqa: E731
    nt = int(np.float32(np.float32(js["TMAX"]) / np.float32(js["DT"])))
    lines = out.splitlines()
    assert lines[0] == "# Parameters setting:"
    assert "* Tmax           = %s (sec)" % g(js["TMAX"]) in lines
    assert "* NX, NY, NZ, NT = %d %d %d %d" % (js["NX"], js["NY"], js["NZ"], nt) in lines
    assert "* Fault_grid     = %d %d %d %d" % tuple(js["Fault_grid"]) in lines
    assert "* PML_N          = %d" % js.get("PML_N", 20) in lines
    ft = js.get("Friction_type", 0)
    assert "# Friction type: %d" % ft in lines
    if ft == 0:
        assert "* mu_s           = %s" % g(js["mu_s"]) in lines
        assert "* Dc             = %s (m)" % g(js["Dc"]) in lines
    assert lines[-1].startswith("* smooth_load_T") or lines[-1].startswith("Media")
