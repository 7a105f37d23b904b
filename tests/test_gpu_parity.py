"""HIP product path vs the FP64 oracle on the same seeded inputs (GPU).

Tolerances are norm-wise relative errors, ||hip - oracle|| / ||oracle|| per
wave component over the interior, stated in each test: FP32 kernels against
an FP64 restatement.
"""
import numpy as np
import pytest

import cases
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

dyn = pytest.importorskip("dynfd3dcuda_amd")

FLAGS8 = [(fx, fy, fz) for fx in (-1, 1) for fy in (-1, 1) for fz in (-1, 1)]


def make_pair(case):
    s = dyn.Solver(case["params"])
    cases.apply_to_solver(case, s)
    s.setup()
    o = Oracle(case["params"])
    cases.apply_to_oracle(case, o)
    o.setup()
    return s, o


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / den if den > 0 else np.linalg.norm(a)


def interior(v, d, ncomp):
    v = v.reshape(ncomp, d["nx"], d["nz"], d["ny"])
    return v[:, 3:-3, 3:-3, 3:-3]


def test_setup_metric_matches_oracle():
    case = cases.volume_case(curved=True)
    s, o = make_pair(case)
    d = o.dims()
    m_h = interior(s.get("M"), d, 13)
    m_o = interior(o.get("M"), d, 13)
    for c in range(13):
        assert relerr(m_h[c], m_o[c]) < 1e-6, c


@pytest.mark.parametrize("flags", FLAGS8)
def test_wave_rhs_curved_free_surface(flags):
    """wave_deriv_cu (wave.cu:8-494) incl. free-surface DzV and traction
    image, curvilinear grid, PML off: tolerance 2e-6 per component."""
    case = cases.volume_case(curved=True, usePML=0)
    s, o = make_pair(case)
    d = o.dims()
    W = cases.random_wavefield(d["nx"], d["ny"], d["nz"], seed=3)
    s.put("W", W)
    o.put("W", W)
    s.wave_rhs(*flags)
    o.wave_deriv(*flags)
    h_h = interior(s.get("hW"), d, 9)
    h_o = interior(o.get("hW"), d, 9)
    for c in range(9):
        e = relerr(h_h[c], h_o[c])
        assert e < 2e-6, (flags, c, e)


@pytest.mark.parametrize("layout", [{}, {"WaveBandRound": 1}, {"WaveTallInterior": 1}], ids=["default", "band_round", "tall"])
@pytest.mark.parametrize("pml,cerjan", [(1, 0), (0, 1), (1, 1)])
def test_wave_steps_absorbing(pml, cerjan, layout):
    """Three full RK4 steps without the fault: stencil + ADE-PML + RK +
    Cerjan against the oracle; tolerance 1e-5 per component.  Three tilings:
    the low j band exactly its width (32-wide tiles), rounded up to whole
    64-wide tiles (WaveBandRound), and 64 x 16 interior tiles in 1024-thread
    blocks (WaveTallInterior), on a grid wide enough (NY 150) for interior
    tiles between the bands."""
    case = cases.volume_case(curved=True, usePML=pml, useCerjan=cerjan, NY=150, **layout)
    s, o = make_pair(case)
    d = o.dims()
    W = cases.random_wavefield(d["nx"], d["ny"], d["nz"], seed=5)
    s.put("W", W)
    o.put("W", W)
    for it in range(3):
        s.step(it)
        o.step(it)
    w_h = interior(s.get("W"), d, 9)
    w_o = interior(o.get("W"), d, 9)
    for c in range(9):
        e = relerr(w_h[c], w_o[c])
        assert e < 1e-5, (c, e)
    if pml:
        for f in ("PMLx1", "PMLy2", "PMLz1"):
            e = relerr(s.get(f), o.get(f))
            assert e < 1e-5, (f, e)


def test_stage_sequence_flags():
    """One step = 4 stages with flags negated each stage (main.cu:440-515)."""
    case = cases.volume_case(curved=False, usePML=0)
    s, o = make_pair(case)
    d = o.dims()
    W = cases.random_wavefield(d["nx"], d["ny"], d["nz"], seed=7)
    s.put("W", W)
    o.put("W", W)
    for irk in range(4):
        s.stage(5, irk)
        o.stage(5, irk, *dyn.stage_flags(5, irk))
    e = relerr(interior(s.get("W"), d, 9), interior(o.get("W"), d, 9))
    assert e < 1e-5, e


@pytest.mark.parametrize("curved", [False, True])
def test_fault_sw_steps(curved):
    """Slip-weakening rupture: split-node fault kernels, trial_sw,
    fault2wave, fault_rk and the slip RK against the oracle.

    After 5 steps every field agrees to 2e-5.  After 40 steps the slip rate
    Vs1/Vs2 (a difference of split-node velocities across a friction switch)
    is allowed 1e-3: the oracle itself moves Vs1 by 6.8e-5 after 40 steps
    when only the initial shear stress is perturbed by 6e-8 relative, so FP32
    rounding on every step cannot be held to 1e-4 there.  Integrated and
    volume quantities stay at 1e-4."""
    case = cases.sw_case(curved=curved)
    s, o = make_pair(case)
    d = o.dims()
    for it in range(5):
        s.step(it)
        o.step(it)
    for name in ("Vs1", "Vs2", "Ts1", "Tn", "slip", "FW"):
        e = relerr(s.get(name), o.get(name))
        assert e < 2e-5, (name, e)
    for it in range(5, 40):
        s.step(it)
        o.step(it)
    for name, tol in (("Vs1", 1e-3), ("Vs2", 1e-3), ("Ts1", 1e-4), ("Tn", 1e-4), ("slip", 1e-4), ("FW", 1e-4)):
        e = relerr(s.get(name), o.get(name))
        assert e < tol, (name, e)
    w_h = interior(s.get("W"), d, 9)
    w_o = interior(o.get("W"), d, 9)
    for c in range(9):
        e = relerr(w_h[c], w_o[c])
        assert e < 1e-4, (c, e)
    # the rupture front reached the same nodes within one step
    t_h, t_o = s.get("init_t0"), o.get("init_t0")
    assert np.array_equal(t_h > -1, t_o > -1)
    assert np.max(np.abs(t_h - t_o)) <= case["params"]["DT"] + 1e-6


@pytest.mark.parametrize("curved", [False, True])
def test_fault_smooth_steps(curved):
    """FaultSmooth (smooth_gauss_volume0/1/2, smooth.cu:13-345, once per step
    after the stages): the surface mirror, the half-volume Gaussian average of
    the split-node tractions and the slip-weighted blend, on a rupturing
    slip-weakening fault.  The smoothed tractions FWs after the first step and
    the split nodes and slip rate after 12 steps are compared with the oracle."""
    case = cases.sw_case(curved=curved, FaultSmooth=1)
    s, o = make_pair(case)
    s.step(0)
    o.step(0)
    for name in ("FWs", "FW"):
        e = relerr(s.get(name), o.get(name))
        assert e < 2e-5, (name, e)
    for it in range(1, 12):
        s.step(it)
        o.step(it)
    for name, tol in (("FW", 1e-4), ("Vs1", 1e-3), ("slip", 1e-4), ("Ts1", 1e-4)):
        e = relerr(s.get(name), o.get(name))
        assert e < tol, (name, e)


def test_fault_rs_steps():
    """Rate-and-state (aging law) with the FP64 Newton solve.

    Before nucleation the fault creeps at ~1e-12 m/s, below the FP32
    rounding floor of the stresses that drive it (6e-8 x 1.2e8 Pa / (rho cs)
    ~ 1e-6 m/s), so the first steps are checked in absolute terms: slip rate
    within 1e-5 m/s, state and tractions to 1e-6 relative.  After nucleation
    (30 steps) slip rate is checked to 1e-3 and slip/state/tractions to 1e-4
    relative."""
    case = cases.rs_case()
    s, o = make_pair(case)
    for it in range(5):
        s.step(it)
        o.step(it)
    for name in ("Vs1", "Vs2"):
        assert np.max(np.abs(s.get(name) - o.get(name))) < 1e-5, name
    for name in ("Ts1", "Tn", "State"):
        e = relerr(s.get(name), o.get(name))
        assert e < 1e-6, (name, e)
    for it in range(5, 30):
        s.step(it)
        o.step(it)
    assert np.max(np.abs(o.get("Vs1"))) > 1e-3  # nucleated
    for name, tol in (("Vs1", 1e-3), ("Ts1", 1e-4), ("Tn", 1e-4), ("slip", 1e-3), ("State", 1e-4)):
        e = relerr(s.get(name), o.get(name))
        assert e < tol, (name, e)


def test_metric_bf16_storage():
    """MetricBF16=1 stores the 13 metric/media words in bfloat16 and computes
    in FP32.  The stored words are the bfloat16 of the FP64 setup values to
    one rounding (|rel| <= 2^-8), and a 40-step slip-weakening rupture on a
    curved grid stays close to the FP64 oracle (measured: slip 6.5e-4, Ts1
    2.5e-5, Tn 2.6e-6; the fault kernels read an FP32 copy of the planes
    around the fault); the measured errors are printed."""
    case = cases.sw_case(curved=True, MetricBF16=1)
    s, o = make_pair(case)
    d = o.dims()
    raw = s.get("M_bf16").astype(np.uint32) << 16
    mb = raw.view(np.float32).astype(np.float64).reshape(13, d["nx"], d["nz"], d["ny"])
    mo = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
    sl = (slice(None), slice(3, -3), slice(3, -3), slice(3, -3))
    rel = np.abs(mb[sl] - mo[sl]) / np.maximum(np.abs(mo[sl]), 1e-30)
    assert rel.max() <= 2.0 ** -8 * 1.0001, rel.max()
    for it in range(40):
        s.step(it)
        o.step(it)
    errs = {n: relerr(s.get(n), o.get(n)) for n in ("slip", "Vs1", "Ts1", "Tn")}
    print("bf16 metric, 40 steps, rel L2 vs FP64 oracle:", errs)
    assert errs["slip"] < 2e-3 and errs["Ts1"] < 1e-4 and errs["Tn"] < 1e-5, errs


def test_metric_bf16_delta_storage():
    """MetricBF16=2 stores the metric words (fields 0-9) as the bfloat16 of
    their deviation from a reference value per field (the nine coefficients
    two to a 32-bit word, exported as 5 int32 fields with field 2p in the
    low half and the media palette entry in the high half of the fifth, the
    Jacobian in a word of its own; the midrange of the field over every 4th
    stored point along each axis, and the last, in FP32, is the reference),
    lambda, mu, rho as FP32 planes, and a palette of the distinct media.
    Decoded, each metric word lies within one bfloat16 rounding of its
    deviation (2^-8 |m - ref|, plus FP32 rounding) of the FP32-stored word,
    the media words equal the FP32-stored ones bit for bit, and so does
    each point's palette entry.  A 40-step slip-weakening rupture on a
    curved grid then lies within 2.7e-6 (slip) and 1.7e-7 (Ts1, Tn) of the
    FP64 oracle, against 6.5e-4 and 2.5e-5 with every word in bfloat16;
    the measured errors are printed."""
    case = cases.sw_case(curved=True, MetricBF16=2)
    s, o = make_pair(case)
    d = o.dims()
    shp = (d["nx"], d["nz"], d["ny"])
    pairs = s.get("M_bf16").view(np.uint32).reshape((5,) + shp)  # field 2p low, 2p+1 high
    m2 = np.stack([pairs & 0xffff, pairs >> 16], axis=1).reshape((10,) + shp)
    entry = m2[9].copy()  # the high half of word 4 holds the media palette entry, the Jacobian has its own word
    m2[9] = s.get("M_jac_bf16").view(np.uint32).reshape(shp) & 0xffff
    t = dyn.Solver(dict(case["params"], MetricBF16=0))
    cases.apply_to_solver(case, t)
    t.setup()
    m0 = t.get("M").reshape((13,) + shp).astype(np.float64)
    t.close()
    ix = [np.r_[np.arange(0, n, 4), n - 1] for n in shp]
    smp = m0[:10][:, ix[0]][:, :, ix[1]][:, :, :, ix[2]].reshape(10, -1)
    ref = s.get("M_ref")  # midrange of the FP64 words, in FP32
    np.testing.assert_allclose(ref, 0.5 * (smp.min(1) + smp.max(1)), rtol=1e-6, atol=1e-9)
    ref = ref.reshape(10, 1, 1, 1)
    dec = np.ascontiguousarray(m2 << 16).view(np.float32) + ref
    err = np.abs(dec.astype(np.float64) - m0[:10])
    bound = 2.0 ** -8 * 1.0001 * np.abs(m0[:10] - ref) + 2.0 ** -21 * np.abs(m0[:10])
    assert np.all(err <= bound), float((err - bound).max())
    plain = (((m0[:10].astype(np.float32).view(np.uint32) + 0x8000) >> 16) << 16).view(np.float32)
    print("max metric error: deviation form", float(err.max()), "plain bfloat16 (about)",
          float(np.abs(plain.astype(np.float64) - m0[:10]).max()))
    media = s.get("M_media").reshape((3,) + shp)
    np.testing.assert_array_equal(media, m0[10:13].astype(np.float32))
    # the two halves of the (bimaterial-form) medium: a palette of two
    # entries, each point's entry holding its own FP32 media words
    pal = s.get("M_palette").reshape(-1, 4)
    assert pal.shape[0] == 2 and entry.max() <= 1
    np.testing.assert_array_equal(np.moveaxis(pal[entry][..., :3], -1, 0), media)
    for it in range(40):
        s.step(it)
        o.step(it)
    errs = {n: relerr(s.get(n), o.get(n)) for n in ("slip", "Vs1", "Ts1", "Tn")}
    print("bf16 metric deviations with exact media, 40 steps, rel L2 vs FP64 oracle:", errs)
    assert errs["slip"] < 1e-5 and errs["Ts1"] < 1e-6 and errs["Tn"] < 1e-6, errs


@pytest.mark.parametrize("kind", ["sw_smooth", "rs"])
def test_metric_bf16_delta_fault_options(kind):
    """The bfloat16-deviation form (MetricBF16=2) through the remaining fault
    kernels that read the metric: fault smoothing (FaultSmooth=1) on the
    curved slip-weakening case, and the rate-and-state solve on a curved
    grid; 40 steps against the FP64 oracle.  Measured: smoothing slip
    3.3e-6; rate-state slip 1.8e-4 (FP32 storage: 1.8e-4 on the same case),
    Ts1 4.2e-7.  The bounds are about three times the measured ones."""
    if kind == "sw_smooth":
        case = cases.sw_case(curved=True, FaultSmooth=1, MetricBF16=2)
    else:
        case = cases.rs_case(curved=True, MetricBF16=2)
    s, o = make_pair(case)
    for it in range(40):
        s.step(it)
        o.step(it)
    errs = {n: relerr(s.get(n), o.get(n)) for n in ("slip", "Vs1", "Ts1", "Tn", "State")}
    print(kind, "bf16 metric deviations, 40 steps, rel L2 vs FP64 oracle:", errs)
    assert np.abs(s.get("Vs1")).max() > 0
    for name, tol in BF16_DELTA_TOL[kind].items():
        assert errs[name] < tol, (name, errs)


BF16_DELTA_TOL = {"sw_smooth": {"slip": 1e-5, "Ts1": 1e-6, "Tn": 1e-6},
                  "rs": {"slip": 5e-4, "Ts1": 2e-6, "Tn": 1e-6, "State": 1e-6}}


def test_media3d_input(tmp_path):
    """Media3D (media.cu:140-204): a heterogeneous medium given on the
    interior grid, as arrays through the C ABI and as the Media3D netCDF file
    (vp, vs, rho with dims NX, NZ, NY, written by scipy), gives the oracle's
    lambda, mu, rho (ghosts mirrored) and the same three RK4 steps: 1e-6 on
    the coefficients, 1e-5 on the wavefield."""
    from scipy.io import netcdf_file
    case = cases.volume_case(curved=True)
    p = case["params"]
    NX, NY, NZ = p["NX"], p["NY"], p["NZ"]
    rng = np.random.default_rng(21)
    ii, kk, jj = np.meshgrid(np.arange(NX), np.arange(NZ), np.arange(NY), indexing="ij")
    vp = 5000.0 + 800.0 * np.sin(ii / 5.0) * np.cos(jj / 7.0) + 300.0 * kk / NZ
    vs = vp / np.sqrt(3.0) * (1.0 + 0.05 * rng.uniform(-1, 1, vp.shape))
    rho = 2500.0 + 200.0 * np.cos(kk / 4.0)
    path = str(tmp_path / "media3d.nc")
    with netcdf_file(path, "w") as f:
        for n, d in (("nx", NX), ("nz", NZ), ("ny", NY)):
            f.createDimension(n, d)
        for n, a in (("vp", vp), ("vs", vs), ("rho", rho)):
            v = f.createVariable(n, "f", ("nx", "nz", "ny"))
            v[:] = a.astype(np.float32)
    o = Oracle(p)
    cases.apply_to_oracle(case, o)
    o.set_media3d(vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32))
    o.setup()
    d = o.dims()
    m_o = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
    W = cases.random_wavefield(d["nx"], d["ny"], d["nz"], seed=6)
    o.put("W", W)
    for it in range(3):
        o.step(it)
    w_o = interior(o.get("W"), d, 9)
    for via_file in (False, True):
        q = dict(p, Media3D=path) if via_file else p
        s = dyn.Solver(q)
        cases.apply_to_solver(case, s)
        if not via_file:
            s.set_media3d(vp, vs, rho)
        s.setup()
        m_h = s.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
        for c in (10, 11, 12):
            assert relerr(m_h[c], m_o[c]) < 1e-6, (via_file, c)
        s.put("W", W)
        for it in range(3):
            s.step(it)
        w_h = interior(s.get("W"), d, 9)
        for c in range(9):
            e = relerr(w_h[c], w_o[c])
            assert e < 1e-5, (via_file, c, e)


def test_fault_rs_slip_law_steps():
    """Rate-and-state slip law with flash weakening (Friction_type 2,
    rk4.cu:902-911): the state rate -V/L (psi - psi_ss(V)) with the flash-
    weakened steady state.  L = 0.1 m (the 0.02 m of the aging-law case makes
    the slip-law state equation too stiff for the RK4 step: the oracle itself
    overflows); after 50 steps the patch is weakening dynamically (slip rate
    ~0.25 m/s) and the bounds are those of the aging-law test."""
    case = cases.rs_case(RS_L=0.1)
    case["params"]["Friction_type"] = 2
    case["stress"]["L"] = np.full_like(case["stress"]["L"], 0.1)
    s, o = make_pair(case)
    for it in range(50):
        s.step(it)
        o.step(it)
    assert np.max(np.abs(o.get("Vs1"))) > 0.1  # nucleated
    for name, tol in (("Vs1", 1e-3), ("Ts1", 1e-4), ("Tn", 1e-4), ("slip", 1e-3), ("State", 1e-4)):
        e = relerr(s.get(name), o.get(name))
        print(name, e)
        assert e < tol, (name, e)


def tp_case(TP_n=21, hy=2.0 ** -11):
    """Rate and state with thermal pressurization (Friction_type 3: the slip
    law with flash weakening for the state, rk4.cu:902-911, plus the pore
    pressure of thermpress.cu in the trial), L = 0.1 m as the slip-law test.
    The hydraulic diffusivity (about 4.9e-4 m^2/s) is exact in FP32, the
    precision of the stress input."""
    case = cases.rs_case(RS_L=0.1, Friction_type=3, TP_n=TP_n)
    case["stress"]["L"] = np.full_like(case["stress"]["L"], 0.1)
    case["stress"]["TP_hy"] = np.full_like(case["stress"]["L"], hy)
    return case


@pytest.mark.parametrize("TP_n", [1, 21])
def test_thermpress_phase(TP_n):
    """cgfdm_thermpress (thermpress.cu:22-233) alone, from the same slip
    rate, friction and normal stress (values exact in FP32) on both sides,
    five calls: the FP64 temperature and pressure rises agree with the
    oracle's to 1e-12 relative (both FP64; the GPU's exp differs in the
    last bits).  TP_n = 1 is the diffusion-free case."""
    case = tp_case(TP_n=TP_n)
    s, o = make_pair(case)
    n2 = case["params"]["NY"] * case["params"]["NZ"]
    for side in (s, o):
        side.put("Vs1", np.full(n2, 2.0 ** -6, np.float32))
        side.put("Vs2", np.full(n2, 2.0 ** -8, np.float32))
        side.put("friction", np.full(n2, 0.5, np.float32))
        side.put("Tn", np.full(n2, -1e8, np.float32))
    for _ in range(5):
        s.thermpress()
        o.thermpress()
    for name in ("TP_T", "TP_P"):
        a, b = s.get(name), o.get(name)
        assert np.abs(b).max() > 0
        e = relerr(a, b)
        assert e < 1e-12, (name, e)


def test_fault_tp_steps():
    """Rate and state with thermal pressurization, 40 steps into the
    rupture (slip rate ~0.4 m/s, pore pressure ~1 MPa): slip rate,
    tractions, slip, state and the temperature and pressure rises against
    the oracle, with the bounds of the rate-state tests."""
    case = tp_case()
    s, o = make_pair(case)
    for it in range(40):
        s.step(it)
        o.step(it)
    assert np.max(np.abs(o.get("Vs1"))) > 0.1  # nucleated
    assert np.max(o.get("TP_P")) > 0
    errs = {}
    for name, tol in (("Vs1", 1e-3), ("Ts1", 1e-4), ("Tn", 1e-4), ("slip", 1e-3), ("State", 1e-4),
                      ("TP_T", 1e-3), ("TP_P", 1e-3)):
        errs[name] = relerr(s.get(name), o.get(name))
        assert errs[name] < tol, (name, errs[name])
    print("thermal pressurization, 40 steps, rel L2 vs oracle:", errs)


def test_thermpress_needs_friction_type_3():
    """cgfdm_thermpress on a slip-weakening fault is a state error, and the
    TP output fields are empty there."""
    case = cases.sw_case()
    s = dyn.Solver(case["params"])
    cases.apply_to_solver(case, s)
    s.setup()
    with pytest.raises(dyn.CgfdmError) as e:
        s.thermpress()
    assert e.value.code == dyn.ESTATE
    with pytest.raises(dyn.CgfdmError):  # no thermal-pressurization state to read
        s.field_info("TP_T")


@pytest.mark.parametrize("media", ["two_half", "layered"])
def test_media_palette_same_result(media):
    """With MetricBF16=2 the stage kernels read lambda, mu, rho from the
    media palette (MetricPalette 1) or from the FP32 planes (0): the same
    words, so the same wavefield and fault fields bit for bit after 10 steps
    of a slip-weakening rupture on a curved grid (a two-half medium, and
    the 3-layer 1-D model), with and without the absorbing layers'
    pipelined band kernel."""
    from dynfd3dcuda_amd import configs
    case = cases.sw_case(curved=True, MetricBF16=2)
    if media == "layered":
        case["media1d"] = configs.three_layer_media()
    runs = []
    for pal in (0, 1):
        s = dyn.Solver(dict(case["params"], MetricPalette=pal))
        cases.apply_to_solver(case, s)
        s.setup()
        npal = s.field_info("M_palette")[0] if pal else 0
        for it in range(10):
            s.step(it)
        s.synchronize()
        runs.append({n: s.get(n) for n in ("W", "Vs1", "Ts1", "Tn", "slip")})
        s.close()
    assert npal == 4 * (3 if media == "layered" else 2)
    assert np.max(np.abs(runs[0]["Vs1"])) > 0
    for n in runs[0]:
        np.testing.assert_array_equal(runs[0][n], runs[1][n], err_msg=n)


def test_metric_constant_words_skipped_same_result():
    """With MetricBF16=2 the stage kernels skip loading the 32-bit metric
    words that are +-0 at every point (a coefficient pair equal to its
    reference values on the whole grid: on this grid, bent in x only with
    Cartesian y and z, the pairs (eta_y, eta_z) and (zeta_x, zeta_y)); a
    skipped word decodes to the same reference values, so 10 steps of a slip-weakening rupture on the
    curved grid give the same fields bit for bit with MetricSkipConstant 0
    and 1."""
    case = cases.sw_case(curved=True, MetricBF16=2)
    runs, masks = [], []
    for skip in (0, 1):
        s = dyn.Solver(dict(case["params"], MetricSkipConstant=skip))
        cases.apply_to_solver(case, s)
        s.setup()
        masks.append(s.metric_skip_mask())
        pairs = s.get("M_bf16").view(np.uint32).reshape(5, -1)
        for it in range(10):
            s.step(it)
        s.synchronize()
        runs.append({n: s.get(n) for n in ("W", "Vs1", "Ts1", "Tn", "slip")})
        s.close()
    assert masks[0] == 0 and masks[1] != 0, masks
    for p in range(5):  # the mask names exactly the words that are +-0 in both halves everywhere
        assert bool((masks[1] >> p) & 1) == (not (pairs[p] & 0x7FFF7FFF).any()), (p, masks[1])
    assert np.max(np.abs(runs[0]["Vs1"])) > 0
    for n in runs[0]:
        np.testing.assert_array_equal(runs[0][n], runs[1][n], err_msg=n)
