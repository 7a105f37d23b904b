"""A stage driven one phase at a time through the C ABI, as the reference's
main loop drives its launchers (main.cu:444-522): cgfdm_rup_sensor, then per
stage cgfdm_fault_trial, cgfdm_wave_stage, cgfdm_fault_deriv and
cgfdm_fault_update, then cgfdm_fault_smooth.

* composed in that order the phases give what cgfdm_step gives, bit for bit
  (slip weakening with smoothing, and rate-and-state);
* at the phase boundaries the fields agree with the FP64 oracle run launcher
  by launcher: the trial tractions and the fault-plane T1 after the trial,
  the split-node rates after the derivatives (FP32 against FP64 after 8
  steps: 1e-4 relative, norm-wise);
* launches can go on a caller's stream (cgfdm_set_stream) with the same
  result;
* the tiling options give the same result, and the ADE-PML work of the
  launches covers every point of every absorbing face once."""
import numpy as np
import pytest

import cases
from test_gpu_parity import make_pair, relerr

pytestmark = pytest.mark.gpu

dyn = pytest.importorskip("dynfd3dcuda_amd")

FIELDS = ("W", "FW", "Vs1", "Vs2", "Ts1", "Ts2", "Tn", "slip", "State", "init_t0", "rup_sensor")


def step_by_phases(s, it, smooth, tp=False):
    s.rup_sensor()
    if tp:
        s.thermpress()
    for irk in range(4):
        s.stage_by_phases(it, irk)
    if smooth:
        s.fault_smooth()


@pytest.mark.parametrize("kind", ["sw_smooth", "rs", "tp"])
def test_phases_compose_to_step(kind):
    if kind == "tp":
        from test_gpu_parity import tp_case
        case = tp_case()
    else:
        case = cases.sw_case(curved=True, FaultSmooth=1) if kind == "sw_smooth" else cases.rs_case()
    a = dyn.Solver(case["params"])
    b = dyn.Solver(case["params"])
    for s in (a, b):
        cases.apply_to_solver(case, s)
        s.setup()
    smooth = kind == "sw_smooth"
    for it in range(12):
        a.step(it)
        step_by_phases(b, it, smooth, tp=kind == "tp")
    a.synchronize()
    b.synchronize()
    assert np.max(np.abs(a.get("Vs1"))) > 0
    for name in FIELDS:
        np.testing.assert_array_equal(a.get(name), b.get(name), err_msg=name)
    if kind == "tp":
        for name in ("TP_T", "TP_P"):
            np.testing.assert_array_equal(a.get(name), b.get(name), err_msg=name)


def test_phase_boundaries_match_oracle():
    case = cases.sw_case(curved=True)
    s, o = make_pair(case)
    nt = 8
    for it in range(nt):
        s.step(it)
        o.step(it)
    fx, fy, fz = s.stage_flags_lib(nt, 0)
    assert (fx, fy, fz) == dyn.stage_flags(nt, 0)
    s.fault_trial(nt, 0)
    o.fault_trial(nt, 0, fx, fy, fz)
    errs = {}
    for name in ("tTn", "tTs1", "tTs2", "T11", "T12", "T13", "FW"):
        errs[name] = relerr(s.get(name), o.get(name))
    s.wave_stage(0, fx, fy, fz)
    s.fault_deriv(nt, 0)
    o.fault_deriv(fx, fy, fz)
    errs["FhW"] = relerr(s.get("FhW"), o.get("FhW"))
    print("phase boundaries, rel L2 vs oracle:", errs)
    assert np.linalg.norm(o.get("tTs1")) > 0 and np.linalg.norm(o.get("FhW")) > 0
    for name, e in errs.items():
        assert e < 1e-4, (name, e)


def hip_runtime():
    """The HIP runtime the solver library runs on (the one already mapped
    into this process), for creating a caller's stream."""
    import ctypes
    for line in open("/proc/self/maps"):
        if "libamdhip64" in line:
            return ctypes.CDLL(line.split()[-1])
    pytest.fail("libamdhip64 not mapped")


def test_caller_stream():
    import ctypes
    case = cases.sw_case(curved=True)
    a = dyn.Solver(case["params"])
    b = dyn.Solver(case["params"])
    for s in (a, b):
        cases.apply_to_solver(case, s)
        s.setup()
    hip = hip_runtime()
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        b.set_stream(st.value)
        for it in range(6):
            a.step(it)
            b.step(it)
        assert hip.hipStreamSynchronize(st) == 0
        b.set_stream(None)
        for it in range(6, 8):
            a.step(it)
            b.step(it)
        a.synchronize()
        b.synchronize()
    finally:
        b.set_stream(None)
        hip.hipStreamDestroy(st)
    for name in ("W", "Vs1", "Ts1", "slip"):
        np.testing.assert_array_equal(a.get(name), b.get(name), err_msg=name)


@pytest.mark.parametrize("extra", [{}, {"WaveBandRound": 1, "WaveTallInterior": 1}], ids=["default", "band_round_tall"])
def test_stage_streams_same_result(extra):
    """StageStreams=1 runs the interior tiles and the free-surface passes on
    side streams, fenced by events: they touch disjoint points, so the
    result is the serial one bit for bit (slip weakening on a curved grid
    with PML and free surface, 10 steps)."""
    case = cases.sw_case(curved=True, NY=150, **extra)
    a = dyn.Solver(case["params"])
    b = dyn.Solver(dict(case["params"], StageStreams=1))
    for s in (a, b):
        cases.apply_to_solver(case, s)
        s.setup()
    for it in range(10):
        a.step(it)
        b.step(it)
    a.synchronize()
    b.synchronize()
    assert np.max(np.abs(a.get("Vs1"))) > 0
    for name in FIELDS:
        np.testing.assert_array_equal(a.get(name), b.get(name), err_msg=name)


@pytest.mark.parametrize("layout,exact,ny", [({"WaveBandRound": 1}, True, 150), ({"WaveTallInterior": 1}, True, 150),
                                             ({"WaveBandRound": 1, "WaveTallInterior": 1}, True, 150),
                                             ({"WaveSurfaceTiles": 0}, False, 150),
                                             ({"WaveInteriorKernel": 0}, True, 150),
                                             ({"WaveInteriorKernel": 2}, True, 150),
                                             ({"WaveInteriorKernel": 3}, True, 150),
                                             ({"WaveInteriorKernel": 4}, True, 150),
                                             ({"WaveBandKernel": 0}, True, 150),
                                             ({"WaveBandKernel": 0, "useCerjan": 1}, True, 150),
                                             ({"WaveBandKernel": 2, "useCerjan": 1}, True, 150),
                                             ({"WaveBandKernel": 2}, True, 150),
                                             ({"PmlAuxFused": 0}, True, 150),
                                             ({"WaveSurfaceKernel": 0}, True, 150),
                                             ({"WaveSurfaceKernel": 0, "useCerjan": 1}, True, 150),
                                             ({"WaveBufferSkew": 0}, True, 150)],
                         ids=["band_round", "tall", "band_round_tall", "surface_tiles_off", "general_interior",
                              "interior_64x6", "interior_64x4", "interior_64x8_double_buffer", "general_band",
                              "general_band_cerjan", "band_64x4_cerjan", "band_64x4", "aux_update_separate", "general_surface",
                              "general_surface_cerjan", "buffers_unskewed"])
def test_tilings_give_the_same_result(layout, exact, ny):
    """The tiling options only move points between the interior, band and
    free-surface launches, or between interior kernels of other tile shapes;
    a point's arithmetic is the same in each (beta = 1 and no damping outside
    the absorbing layers).  10 steps of a slip-weakening rupture on a curved
    grid with PML and free surface:
    - moves between the interior and band kernels, the general, 64 x 6,
      64 x 4 and double-buffered 64 x 8 interior kernels, the general
      kernel and 64 x 4 tiles on the absorbing bands (64 x 8 on this grid
      by default)
      (with ADE-PML, and with PML and the Cerjan sponge together), the
      ADE-PML auxiliary update of every band point in its own kernel
      instead of in the band tiles (which by default advance the points on
      one face only and leave the edges of the absorbing box to that
      kernel; NX 30 puts 16 of the 30 planes in the x bands, so most y- and
      z-band points are such edges), and the general kernel instead of the
      pipelined one on the free-surface tiles, and wavefield buffers at
      their allocations' bases instead of skewed, give the default tiling's
      result bit for bit;
    - WaveSurfaceTiles 0 moves rows between the free-surface kernel and the
      others, whose code the compiler contracts differently (multiply-adds
      fused or not), so the result agrees to FP32 rounding: 1e-5 relative
      per field after 10 steps (both stay 4e-6 from the FP64 oracle)."""
    case = cases.sw_case(curved=True, NY=ny, **{k: v for k, v in layout.items() if k == "useCerjan"})
    a = dyn.Solver(case["params"])
    b = dyn.Solver(dict(case["params"], **layout))
    for s in (a, b):
        cases.apply_to_solver(case, s)
        s.setup()
    for it in range(10):
        a.step(it)
        b.step(it)
    a.synchronize()
    b.synchronize()
    assert np.max(np.abs(a.get("Vs1"))) > 0
    # the ADE-PML auxiliary fields of the faces that absorb (the top is free)
    for name in FIELDS + ("PMLx1", "PMLx2", "PMLy1", "PMLy2", "PMLz1"):
        x, y = a.get(name), b.get(name)
        if exact:
            np.testing.assert_array_equal(x, y, err_msg=name)
        else:
            x, y = x.astype(np.float64), y.astype(np.float64)
            den = np.linalg.norm(y)
            e = np.linalg.norm(x - y) / den if den > 0 else np.linalg.norm(x)
            assert e < 1e-5, (name, e)


@pytest.mark.parametrize("layout", [{}, {"PmlAuxFused": 0}, {"WaveBandKernel": 0}, {"WaveBandKernel": 2},
                                    {"WaveSurfaceTiles": 0}, {"FreeSurface": 0}],
                         ids=["default", "aux_separate", "general_band", "band_64x4", "surface_tiles_off",
                              "no_free_surface"])
def test_aux_pairs_cover_every_face_point_once(layout):
    """The ADE-PML work the launches report (cgfdm_profile_aux) partitions
    the absorbing faces: every (point, face) pair has its auxiliary value
    read by exactly one stage-kernel launch and its auxiliary ODE step taken
    by exactly one launch (the band tiles for points on one face with the
    fused update, pml_aux_kernel for the edges of the absorbing box, the
    free-surface rows, or everything).  A box table of pml_aux_kernel that
    missed or repeated a box shows here as a wrong total."""
    case = cases.sw_case(curved=True, NY=150, **layout)
    s = dyn.Solver(case["params"])
    cases.apply_to_solver(case, s)
    s.setup()
    d = s.dims()
    p = case["params"]
    N, n = p["PML_N"], (d["ni"], d["nj"], d["nk"])
    on = [1, 1, 1, 1, 1, 0 if p.get("FreeSurface", 0) else 1]
    total = sum(N * n[(f // 2 + 1) % 3] * n[(f // 2 + 2) % 3] for f in range(6) if on[f])
    aux = s.profile_aux()
    s.close()
    assert total > 0
    assert sum(rd for rd, _ in aux.values()) == total, aux
    assert sum(up for _, up in aux.values()) == total, aux
    assert aux["pml_aux"][0] == 0
    fused = layout.get("PmlAuxFused", 1) and layout.get("WaveBandKernel", 2) >= 1
    assert (aux["stage_band"][1] > 0) == bool(fused), aux
