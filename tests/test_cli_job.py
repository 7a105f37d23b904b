"""The drop-in path end to end: a job directory in the reference's layout
(params.json naming Fault_geometry.nc and Fault_init_stress.nc, OUT/), run by
the command-line driver `dynfd3dcuda_amd/cgfdm params.json` as the reference's
`a.out params.json` is (main.cu:123-580), must write the reference's output
files (fault_mpi000000.nc, wave_{xy,xz,yz}_mpi000000.nc; io_fault.cu:16-133,
io_wave_*.cu) with its variable schema, and the same values as the same job
driven through the C ABI in process.  The netCDF inputs are written with
scipy (CDF-1), an implementation independent of the solver's reader."""
import json
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

import cases

pytestmark = pytest.mark.gpu

dyn = pytest.importorskip("dynfd3dcuda_amd")

FAULT_VARS_2D = ["x", "y", "z", "init_t0", "C0", "str_peak", "united", "a", "b", "Vw"]
FAULT_VARS_T = ["Vs1", "Vs2", "tn", "ts1", "ts2", "Us0", "State", "rup_index_y", "rup_index_z", "rup_sensor",
                "TP_T", "TP_P"]


def write_job(d, out_dir, nt=24, case=None):
    case = case or cases.sw_case(curved=True)
    p = dict(case["params"])
    NY, NZ = p["NY"], p["NZ"]
    with netcdf_file(os.path.join(d, "Fault_geometry.nc"), "w") as f:
        f.createDimension("ny", NY)
        f.createDimension("nz", NZ)
        geom = case.get("geom") or cases.bent_geometry(NY, NZ, p["DH"], amp=0.0)
        for k, v in zip("xyz", geom):
            var = f.createVariable(k, "f", ("nz", "ny"))
            var[:] = np.asarray(v, np.float32)
    with netcdf_file(os.path.join(d, "Fault_init_stress.nc"), "w") as f:
        f.createDimension("ny", NY)
        f.createDimension("nz", NZ)
        for k, v in case["stress"].items():
            var = f.createVariable(k, "f", ("nz", "ny"))
            var[:] = np.asarray(v, np.float32).reshape(NZ, NY)
    p.update({"INPORT_GRID_TYPE": 1, "Fault_geometry": "Fault_geometry.nc", "Fault_init_stress": "Fault_init_stress.nc",
              "OUT": out_dir, "TMAX": nt * p["DT"] + 0.5 * p["DT"], "EXPORT_TIME_SKIP": 4,
              "EXPORT_WAVE_SLICE_X": p["NX"] // 2, "EXPORT_WAVE_SLICE_Y": NY // 2, "EXPORT_WAVE_SLICE_Z": NZ - 1})
    with open(os.path.join(d, "params.json"), "w") as f:
        json.dump(p, f, indent=1)
    return p


def read_nc(path):
    with netcdf_file(path, "r", mmap=False) as f:
        return {k: v[:].copy() for k, v in f.variables.items()}, dict(f.dimensions)


@pytest.mark.parametrize("metric_bf16", [0, 2])
def test_cli_job_directory(tmp_path, metric_bf16):
    """The job as the reference runs it, with the metric in FP32 and in the
    bfloat16-deviation form (MetricBF16=2)."""
    job = str(tmp_path)
    p = write_job(job, "OUT", case=cases.sw_case(curved=True, MetricBF16=metric_bf16))
    r = subprocess.run([dyn.CLI_PATH, "params.json"], cwd=job, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "> Total" in r.stdout and ">      0, total" in r.stdout
    nt = int(p["TMAX"] / p["DT"])
    fv, fd = read_nc(os.path.join(job, "OUT", "fault_mpi000000.nc"))
    for k in FAULT_VARS_2D + FAULT_VARS_T:
        assert k in fv, k
    nrec = (nt + p["EXPORT_TIME_SKIP"] - 1) // p["EXPORT_TIME_SKIP"]
    assert fv["Vs1"].shape[0] == nrec and fv["Vs1"].shape[1:] == fv["x"].shape
    for name in ("xy", "xz", "yz"):
        wv, _ = read_nc(os.path.join(job, "OUT", f"wave_{name}_mpi000000.nc"))
        for k in ("Vx", "Vy", "Vz"):
            assert wv[k].shape[0] == nrec, (name, k)
    # the same job through the C ABI in process (outputs into OUT2)
    q = json.load(open(os.path.join(job, "params.json")))
    q["OUT"] = "OUT2"
    with open(os.path.join(job, "params2.json"), "w") as f:
        json.dump(q, f)
    cwd = os.getcwd()
    os.chdir(job)
    try:
        s = dyn.Solver(path="params2.json")
        s.setup()
        s.open_outputs()
        for it in range(nt):
            s.write_outputs(it)
            s.step(it)
        s.synchronize()
        s.close_outputs()
        s.close()
    finally:
        os.chdir(cwd)
    gv, _ = read_nc(os.path.join(job, "OUT2", "fault_mpi000000.nc"))
    for k in FAULT_VARS_2D + FAULT_VARS_T:
        np.testing.assert_array_equal(fv[k], gv[k], err_msg=k)
    assert np.abs(fv["Vs1"][-1]).max() > 0  # the rupture moved



def test_cli_thermal_pressurization_job(tmp_path):
    """A rate-and-state job with thermal pressurization (Friction_type 3,
    TP_n, TP_hy in Fault_init_stress.nc) through the command-line driver:
    the fault file carries non-zero TP_T and TP_P records (the rises on the
    fault, io_fault.cu:262-266), equal to what the same job gives through
    the C ABI."""
    from test_gpu_parity import tp_case
    job = str(tmp_path)
    p = write_job(job, "OUT", nt=40, case=tp_case())
    r = subprocess.run([dyn.CLI_PATH, "params.json"], cwd=job, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    fv, _ = read_nc(os.path.join(job, "OUT", "fault_mpi000000.nc"))
    assert np.abs(fv["TP_P"][-1]).max() > 0 and np.abs(fv["TP_T"][-1]).max() > 0
    q = json.load(open(os.path.join(job, "params.json")))
    q["OUT"] = "OUT2"
    with open(os.path.join(job, "params2.json"), "w") as f:
        json.dump(q, f)
    cwd = os.getcwd()
    os.chdir(job)
    try:
        s = dyn.Solver(path="params2.json")
        s.setup()
        s.open_outputs()
        s.run(0, int(p["TMAX"] / p["DT"]), write_outputs=True)
        s.close_outputs()
        s.close()
    finally:
        os.chdir(cwd)
    gv, _ = read_nc(os.path.join(job, "OUT2", "fault_mpi000000.nc"))
    for k in ("TP_T", "TP_P", "Vs1", "State", "Us0"):
        np.testing.assert_array_equal(fv[k], gv[k], err_msg=k)
