"""Each benchmark workload's own configuration on the GPU (HIP path through
the C ABI) against the FP64 oracle, at full size.

* flat_sw_100x200x100: plain flat grid, slip weakening, Cerjan sponge (the
  rupture path that damps after fault2wave, main.cu:507-509 and
  cerjan.cu:166-200): 40 steps, the rupture under way;
* curved_sw_400x300x200: curved grid, topography, bent fault, 3-layer medium,
  20-layer ADE-PML: 5 steps;
* curved_rs_400x300x200 (reference Newton rule) and curved_rs_mlp_400x300x200
  (learned guess, the committed network): 3 steps of the creeping fault
  before the nucleation ramp, compared in absolute terms below the FP32
  rounding floor of the tractions that drive it;
* large_rs_mlp_bf16_1600x1000x500 (800 M points, 155 GB of HBM): 3 steps on
  the full grid checked by properties (no NaN/Inf anywhere, ghost layers
  zero, fault fields finite, slip never decreasing, Newton without
  fallbacks), and the same physics on the 100x1000x500 sub-slab around the
  fault against the oracle (the full grid needs ~300 GB of host memory in
  FP64), from a smooth seeded initial wavefield so that the stencil, the
  bfloat16-deviation metric and the fault all carry signal.

Every output of the fault file is compared (io_fault.cu:37-67): Vs1, Vs2,
Tn, Ts1, Ts2, slip, State, init_t0, rup_sensor, plus the split nodes FW and
the volume W.  Tolerances are written per check; the measured errors are
printed.
"""
import os

import numpy as np
import pytest

from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

dyn = pytest.importorskip("dynfd3dcuda_amd")
from dynfd3dcuda_amd import configs  # noqa: E402

NET = os.path.join(os.path.dirname(dyn.__file__), "data", "newton_guess_4x64x64x1.f32")


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / den if den > 0 else np.linalg.norm(a)


def f32(a):
    """The input as the solver receives it: the reference reads its input
    files as float (coord.cu:84-146, init.cu:923-1003), and so does the C
    ABI."""
    return np.asarray(a, np.float32).astype(np.float64)


def oracle_for(case):
    """The oracle on the same inputs the GPU solver receives, float32-rounded
    (a topography height rounded differently can move a grid point across a
    layer boundary of the 1-D medium, and with it that point's media)."""
    o = Oracle(case["params"])
    if "geom" in case:
        o.set_geometry(*[f32(a) for a in case["geom"]])
    if "media1d" in case:
        o.set_media1d(*[f32(a) for a in case["media1d"]])
    if case.get("topo") is not None:
        o.set_topography(f32(case["topo"]))
    for k, v in case.get("stress", {}).items():
        o.set_stress(k, f32(v))
    o.setup()
    return o


def solver_for(case):
    s = dyn.Solver(case["params"])
    configs.apply(case, s)
    s.setup()
    return s


def wave_interior(v, d):
    return v.reshape(9, d["nx"], d["nz"], d["ny"])[:, 3:-3, 3:-3, 3:-3]


def compare_sw(s, o, d, DT, tol_fault, tol_rate, tol_wave):
    """Slip-weakening outputs after a run: relative L2 per field, Ts2 against
    the shear scale (it is near zero on a strike-slip fault), init_t0 the
    same ruptured set within one step.  Each wavefield component is measured
    against the largest component norm of its kind (velocities, stresses):
    on a strike-slip rupture Vz, Tzz and Txz are an order of magnitude
    smaller than Vy and Txy, and their own norms would weigh the FP32
    rounding of the larger components that feed them."""
    errs = {}
    for name in ("Ts1", "Tn", "slip", "FW"):
        errs[name] = relerr(s.get(name), o.get(name))
    for name in ("Vs1", "Vs2"):
        errs[name] = relerr(s.get(name), o.get(name))
    ts1 = np.abs(o.get("Ts1")).max()
    errs["Ts2_abs_over_Ts1max"] = np.abs(s.get("Ts2") - o.get("Ts2")).max() / ts1
    errs["rup_sensor"] = relerr(s.get("rup_sensor"), o.get("rup_sensor"))
    w_h, w_o = wave_interior(s.get("W"), d), wave_interior(o.get("W"), d)
    nrm = [np.linalg.norm(w_o[c]) for c in range(9)]
    for c in range(9):
        kind = max(nrm[0:3]) if c < 3 else max(nrm[3:9])
        errs[f"W{c}"] = np.linalg.norm(w_h[c].astype(np.float64) - w_o[c]) / kind
    print({k: f"{v:.2e}" for k, v in errs.items()})
    for name in ("Ts1", "Tn", "slip", "FW"):
        assert errs[name] < tol_fault, (name, errs)
    for name in ("Vs1", "Vs2"):
        assert errs[name] < tol_rate, (name, errs)
    assert errs["Ts2_abs_over_Ts1max"] < tol_fault, errs
    assert errs["rup_sensor"] < 2e-3, errs
    for c in range(9):
        assert errs[f"W{c}"] < tol_wave, (c, errs)
    t_h, t_o = s.get("init_t0"), o.get("init_t0")
    assert np.any(t_o > -1), "the rupture has started"
    assert np.array_equal(t_h > -1, t_o > -1)
    assert np.max(np.abs(t_h - t_o)[t_o > -1]) <= DT + 1e-6


def test_flat_sw_100x200x100_full_size():
    """Plain method on its own grid: 40 steps, the Cerjan sponge damping the
    rupture's stage outputs.  Bounds of the 40-step rupture tests (slip rate
    1e-3, the rest 1e-4)."""
    case = configs.flat_sw_100x200x100(nt=60)
    p = case["params"]
    assert p["useCerjan"] == 1 and p["Rupture"] == 1 and p["usePML"] == 0
    s, o = solver_for(case), oracle_for(case)
    d = o.dims()
    for it in range(40):
        s.step(it)
        o.step(it)
    compare_sw(s, o, d, p["DT"], tol_fault=1e-4, tol_rate=1e-3, tol_wave=1e-4)


def test_curved_sw_400x300x200_full_size():
    """Curved grid with topography, bent fault, 3 layers, 20-layer PML: 5
    steps; bounds of the 5-step tests (2e-5) for the fault and the volume.
    The metric and media the GPU computed at setup are compared first."""
    case = configs.curved_sw_400x300x200(nt=20)
    s, o = solver_for(case), oracle_for(case)
    d = o.dims()
    m_h = s.get("M").reshape(13, d["nx"], d["nz"], d["ny"])[:, 3:-3, 3:-3, 3:-3]
    m_o = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])[:, 3:-3, 3:-3, 3:-3]
    # largest error of each field relative to the field's largest magnitude
    # (fields that vanish on this grid, e.g. xi_z, against the row's scale)
    scale = [max(np.abs(m_o[c]).max(), np.abs(m_o[3 * (c // 3):3 * (c // 3) + 3]).max() if c < 9 else 0.0) for c in range(13)]
    merr = [float(np.abs(m_h[c] - m_o[c]).max() / scale[c]) for c in range(13)]
    del m_h, m_o
    print("metric/media max rel", ["%.1e" % e for e in merr])
    assert max(merr) < 1e-6, merr
    for it in range(5):
        s.step(it)
        o.step(it)
    compare_sw(s, o, d, case["params"]["DT"], tol_fault=2e-5, tol_rate=2e-5, tol_wave=2e-5)
    for f in ("PMLx1", "PMLy2", "PMLz1"):
        e = relerr(s.get(f), o.get(f))
        assert e < 1e-5, (f, e)


@pytest.mark.parametrize("name", ["curved_rs_400x300x200", "curved_rs_mlp_400x300x200"])
def test_curved_rs_400x300x200_full_size(name):
    """Rate and state on the curved grid, 3 steps before the nucleation ramp
    has grown: the fault creeps at ~1e-12 m/s against an FP32 traction
    rounding floor of ~1e-6 m/s, so slip rates are compared in absolute terms
    (1e-5 m/s) and the volume against the traction scale (1e-6 of 120 MPa,
    1e-5 m/s); state and tractions relative.  The rupture sensor is the
    square of a difference of second differences of Ts1 (rup_sensor.cu:
    37-80), taken in FP32 as the reference does: on the creeping fault those
    differences are ~1e3 Pa of a ~7.5e7 Pa word, so FP32 rounding alone moves
    the sensor by about 2e-2 relative; it is held to 5e-2 here (on the
    rupturing slip-weakening workloads, where the differences are large, to
    2e-3).  The learned guess (NewtonMode 2) runs against the oracle's
    reference Newton rule."""
    if name.endswith("mlp_400x300x200"):
        case = configs.curved_rs_mlp_400x300x200(nt=20, mlp_path=NET)
    else:
        case = configs.curved_rs_400x300x200(nt=20)
    s, o = solver_for(case), oracle_for(case)
    d = o.dims()
    for it in range(3):
        s.step(it)
        o.step(it)
    errs = {n: relerr(s.get(n), o.get(n)) for n in ("Ts1", "Tn", "State", "rup_sensor")}
    errs.update({n + "_abs": np.abs(s.get(n) - o.get(n)).max() for n in ("Vs1", "Vs2", "slip")})
    errs["Ts2_abs_over_Ts1max"] = np.abs(s.get("Ts2") - o.get("Ts2")).max() / np.abs(o.get("Ts1")).max()
    w_h, w_o = wave_interior(s.get("W"), d), wave_interior(o.get("W"), d)
    errs["W_vel_abs"] = np.abs(w_h[:3] - w_o[:3]).max()
    errs["W_str_abs"] = np.abs(w_h[3:] - w_o[3:]).max()
    print(name, {k: f"{v:.2e}" for k, v in errs.items()})
    for n in ("Ts1", "Tn", "State"):
        assert errs[n] < 1e-6, (n, errs)
    assert errs["rup_sensor"] < 5e-2, errs
    assert errs["Ts2_abs_over_Ts1max"] < 1e-6, errs
    for n in ("Vs1_abs", "Vs2_abs", "W_vel_abs"):
        assert errs[n] < 1e-5, (n, errs)
    assert errs["slip_abs"] < 1e-5 * 3 * case["params"]["DT"] * 4, errs
    assert errs["W_str_abs"] < 1e-6 * 120e6, errs
    if case["params"].get("NewtonMode") == 2:
        st = s.newton_stats()
        assert st["solves"] > 0 and st["fallbacks"] == 0, st


def test_large_rs_mlp_bf16_1600x1000x500_full_size_properties():
    """The 800 M-point workload as configured (metric as bfloat16 deviations,
    FP32 media, learned Newton guess): 3 steps, then on the device the whole
    wavefield has no NaN/Inf and zero ghost layers; the fault fields are
    finite, slip does not decrease, the friction coefficient stays in (0, 1),
    and no Newton solve fell back."""
    case = configs.large_rs_mlp_bf16_1600x1000x500(nt=10, mlp_path=NET)
    s = solver_for(case)
    d = s.dims()
    assert d["ni"] * d["nj"] * d["nk"] == 800_000_000
    slip_prev = None
    for it in range(3):
        s.step(it)
        slip = s.get("slip")
        if slip_prev is not None:
            assert np.all(slip >= slip_prev), it
        slip_prev = slip
    chk = s.check_field("W")
    print("W check", chk)
    assert chk["nonfinite"] == 0 and chk["ghost_nonzero"] == 0, chk
    for n in ("FW", "Vs1", "Vs2", "Ts1", "Ts2", "Tn", "rup_sensor"):
        c = s.check_field(n)
        assert c["nonfinite"] == 0, (n, c)
    for n in ("slip", "State"):
        assert np.all(np.isfinite(s.get(n))), n
    live = s.get("united") == 0
    fr = s.get("friction")[live]
    assert np.all((fr > 0) & (fr < 1)), (fr.min(), fr.max())
    st = s.newton_stats()
    assert st["solves"] > 0 and st["fallbacks"] == 0, st
    s.close()


def smooth_wavefield(d, DH, seed):
    """A smooth seeded wavefield on the padded grid (sum of three plane waves
    per component; velocities ~0.1 m/s, stresses ~1 MPa), zero ghosts."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = d["nx"], d["ny"], d["nz"]
    i = np.arange(nx)[:, None, None]
    k = np.arange(nz)[None, :, None]
    j = np.arange(ny)[None, None, :]
    W = np.zeros((9, nx, nz, ny), np.float32)
    for c in range(9):
        amp = 0.1 if c < 3 else 1e6
        acc = np.zeros((nx, nz, ny), np.float32)
        for _ in range(3):
            kv = rng.uniform(-1, 1, 3) * 2 * np.pi / 40.0  # wavelengths of ~40 cells and more
            ph = rng.uniform(0, 2 * np.pi)
            acc += np.sin(kv[0] * i + kv[1] * k + kv[2] * j + ph).astype(np.float32)
        W[c] = amp / 3 * acc
    W[:, :3] = W[:, -3:] = 0
    W[:, :, :3] = W[:, :, -3:] = 0
    W[:, :, :, :3] = W[:, :, :, -3:] = 0
    return W


def test_large_rs_mlp_bf16_subslab_matches_oracle():
    """The large workload's physics (25 m grid, bent fault, topography rows,
    3 layers, rate and state with the learned guess, bfloat16-deviation
    metric) on its 100x1000x500 sub-slab around the fault against the FP64
    oracle, 2 steps from a smooth seeded wavefield.  The bfloat16 metric
    words carry a relative rounding of 2^-8 of each word's deviation from
    its field's reference, so the volume is held to 2e-4, the fault
    tractions and state to 1e-5, the split nodes to 5e-4 and the slip rates
    (differences of split-node velocities; Vs2, along dip, is the smaller) to
    1e-3 (measured: W 1.5-5.4e-5, Ts1 1.1e-6, FW 1.4e-4, Vs2 2.8e-4)."""
    case = configs.large_rs_mlp_bf16_1600x1000x500(sub_nx=100, nt=10, mlp_path=NET)
    s, o = solver_for(case), oracle_for(case)
    d = o.dims()
    W = smooth_wavefield(d, case["params"]["DH"], seed=11)
    s.put("W", W)
    o.put("W", W)
    del W
    for it in range(2):
        s.step(it)
        o.step(it)
    errs = {n: relerr(s.get(n), o.get(n)) for n in ("Ts1", "Tn", "State", "FW", "Vs1", "Vs2")}
    w_h, w_o = wave_interior(s.get("W"), d), wave_interior(o.get("W"), d)
    for c in range(9):
        errs[f"W{c}"] = relerr(w_h[c], w_o[c])
    del w_h, w_o
    print({k: f"{v:.2e}" for k, v in errs.items()})
    for c in range(9):
        assert errs[f"W{c}"] < 2e-4, (c, errs)
    for n in ("Ts1", "Tn", "State"):
        assert errs[n] < 1e-5, (n, errs)
    assert errs["FW"] < 5e-4, errs
    for n in ("Vs1", "Vs2"):
        assert errs[n] < 1e-3, (n, errs)


def test_nonfinite_output_record_is_reported(tmp_path):
    """A fault record holding NaN makes cgfdm_write_outputs return
    CGFDM_ENONFINITE after writing it, and the device check counts NaN in
    the wavefield."""
    case = configs.flat_sw_100x200x100(nt=10)
    case["params"].update({"OUT": str(tmp_path / "out"), "EXPORT_TIME_SKIP": 1})
    s = solver_for(case)
    s.open_outputs()
    s.write_outputs(0)
    vs = s.get("Vs1")
    vs[123] = np.nan
    s.put("Vs1", vs)
    with pytest.raises(dyn.CgfdmError) as e:
        s.write_outputs(1)
    assert e.value.code == dyn.ENONFINITE
    assert "Vs1" in str(e.value)
    d = s.dims()
    W = s.get("W").reshape(9, d["nx"], d["nz"], d["ny"])
    W[4, 10, 10, 10] = np.inf  # an interior word
    s.put("W", W)
    chk = s.check_field("W")
    assert chk["nonfinite"] == 1, chk
    s.close_outputs()
    s.close()
