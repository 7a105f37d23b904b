"""netCDF classic-format reader/writer of the solver (csrc/netcdf_classic.cpp).

Checked against scipy.io.netcdf_file, an independent implementation of the
same published format: a file our writer makes (fixed and record variables,
float/int/double, an unlimited dimension: the fault_mpi*.nc / wave_*.nc
schema of io_fault.cu:16-133) must read back identically in scipy, and files
scipy writes in CDF-1 and CDF-2 (the Fault_geometry.nc / Fault_init_stress.nc
inputs) must read back identically through our reader.
"""
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dynfd3dcuda_amd", "csrc")


@pytest.fixture(scope="module")
def tool(tmp_path_factory):
    d = tmp_path_factory.mktemp("nc")
    exe = str(d / "nc_tool")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "nc_tool.cpp"),
                           os.path.join(CSRC, "netcdf_classic.cpp"), "-o", exe])
    return exe


def read_with_tool(tool, path, var):
    out = subprocess.check_output([tool, "read", path, var]).decode().split()
    return np.array([float(x) for x in out])


def test_our_writer_read_by_scipy(tool, tmp_path):
    p = str(tmp_path / "fault.nc")
    subprocess.check_call([tool, "write", p])
    with netcdf_file(p, "r", mmap=False) as f:
        assert f.dimensions["nt"] is None and f.dimensions["nz"] == 3 and f.dimensions["ny"] == 4
        x = f.variables["x"][:].copy()
        u = f.variables["united"][:].copy()
        v = f.variables["Vs1"][:].copy()
        t = f.variables["t"][:].copy()
        assert f.variables["x"].typecode() == "f" and f.variables["united"].typecode() == "i"
    q = np.arange(12)
    np.testing.assert_array_equal(x, (0.5 * q - 1.0).reshape(3, 4).astype(np.float32))
    np.testing.assert_array_equal(u, (q % 3).reshape(3, 4))
    assert v.shape == (3, 3, 4)
    for r in range(3):
        np.testing.assert_array_equal(v[r], (100.0 * r + q).reshape(3, 4))
    np.testing.assert_array_equal(t, [0.0, 0.25, 0.5])


@pytest.mark.parametrize("version", [1, 2])
def test_scipy_writer_read_by_ours(tool, tmp_path, version):
    p = str(tmp_path / f"stress_v{version}.nc")
    rng = np.random.default_rng(version)
    a = rng.standard_normal((5, 7)).astype(np.float32)
    b = rng.standard_normal((5, 7))
    c = rng.integers(-5, 5, (5, 7)).astype(np.int32)
    s = rng.integers(-300, 300, (5, 7)).astype(np.int16)
    with netcdf_file(p, "w", version=version) as f:
        f.history = "test"
        f.createDimension("nz", 5)
        f.createDimension("ny", 7)
        for name, arr, tc in (("Tx", a, "f"), ("Ty", b, "d"), ("rup", c, "i"), ("sh", s, "h")):
            v = f.createVariable(name, tc, ("nz", "ny"))
            v.units = "Pa"
            v[:] = arr
    np.testing.assert_array_equal(read_with_tool(tool, p, "Tx"), a.ravel().astype(np.float64))
    np.testing.assert_array_equal(read_with_tool(tool, p, "Ty"), b.ravel())
    np.testing.assert_array_equal(read_with_tool(tool, p, "rup"), c.ravel())
    np.testing.assert_array_equal(read_with_tool(tool, p, "sh"), s.ravel())
