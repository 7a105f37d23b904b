"""Pins of the oracle's fault-side, free-surface and friction functions
against mathematics (CPU), each shown to fail on planted mistakes.

The reference ships no golden vectors for these routines, so each one is
pinned by something the mathematics fixes:

* a wavefield that is linear in space is an exact solution of the discrete
  scheme: every operator of this path (the DRP/MacCormack and the one-sided
  split-node operators, the 2-4 and 2-2 operators next to the surface, the
  free-surface image) differentiates linear fields exactly, and the elastic
  equations then make the solution linear in time, which RK4 integrates
  exactly.  A locked fault (stress far below strength) keeps the medium
  continuous, so the split-node machinery must reproduce the same solution:
  trial traction = the exact traction (trial_slipweakening.cu:363-378,
  trial_ratestate.cu:453-716 with a vanishing slip rate), split-node
  velocity and stress rates (fault_dvelo.cu:17-389, fault_dstrs_f.cu /
  fault_dstrs_b.cu), the fault RK and the RK-weighted stage tractions
  (rk4.cu:466-789), fault2wave (transform.cu:224-403), and the free-surface
  rows (wave.cu:205-246, 398-487) for a traction-free field;
* closed forms of the state laws at constant slip rate (rk4.cu:791-938);
* the closed-form RK4 solution of the ADE-PML auxiliary ODE
  (pmlx.cu:126-210) and the traction-free form of its surface correction;
* the rupture sensor of a quadratic shear traction (rup_sensor.cu:7-97);
* the free-surface split-node maps (fault_coef.cu:855-1010): traction-free
  on each side and the minus-to-plus map inverse to the plus-to-minus one.

Each pin is run again with a numbered mistake planted in the oracle
(oracle.binding.set_mutation) and must then fail.
"""
import math

import numpy as np
import pytest

import cases
from oracle.binding import Oracle, set_mutation


def _rotation(a, b, c):
    ca, sa, cb, sb, cc, sc = math.cos(a), math.sin(a), math.cos(b), math.sin(b), math.cos(c), math.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def _affine_coords(p, Amap):
    """x = Amap (i, j, k) DH on the padded grid, layout [3][nx][nz][ny]."""
    nx, ny, nz = p["NX"] + 6, p["NY"] + 6, p["NZ"] + 6
    I, K, J = np.meshgrid(np.arange(nx), np.arange(nz), np.arange(ny), indexing="ij")
    idx = np.stack([I, J, K]).astype(np.float64) * p["DH"]
    return np.einsum("ab,bijk->aijk", Amap, idx)


def _voigt(S):
    """[..., 3, 3] symmetric tensor -> the wavefield's 6 stress words."""
    return np.stack([S[..., 0, 0], S[..., 1, 1], S[..., 2, 2], S[..., 0, 1], S[..., 0, 2], S[..., 1, 2]])


def _hooke(eps, lam, mu):
    return lam * np.trace(eps) * np.eye(3) + 2 * mu * eps


@pytest.fixture(autouse=True)
def _no_mutation():
    set_mutation(0)
    yield
    set_mutation(0)


# ---------------------------------------------------------------------------
# Linear exact solution through a locked fault

# The split nodes read the never-updated split-node words of the 30-node
# 'united' border (init.cu:806-816), and every stage carries that error three
# nodes further (the stencils' reach), as it carries the zero ghosts' error in
# from the grid edges: after N_STEPS steps only nodes 12 N_STEPS from both are
# clean, so the grid leaves a clean core of split nodes under the surface.
N_STEPS = 2
REACH = 12 * N_STEPS
NX_L, NY_L, NZ_L = 2 * REACH + 24, 60 + 2 * REACH + 6, 30 + REACH + 10


def _linear_problem(friction, Amap):
    """A locked fault on the x-mid plane of an affine grid x = Amap (i,j,k) DH
    with the free surface on top, homogeneous medium, and a linear wavefield
    whose stress is traction-free on the surface (and, for rate and state,
    whose normal traction on the fault and traction rate on the fault vanish,
    as the rate-state trial keeps the x traction at its initial value,
    trial_ratestate.cu:797, and its DxV_T1 term reads no traction rate,
    init.cu:835-847).  The split nodes' DxV_T1 term reads the traction rate
    only inside Fault_grid less three nodes (init.cu:820-828), so the fault
    grid reaches three rows past the surface.  Returns the oracle after setup
    and the exact solution's pieces."""
    p = cases.base_params(NX=NX_L, NY=NY_L, NZ=NZ_L, usePML=0, FreeSurface=1, NormalRefine=0, Rupture=1,
                          Friction_type=friction, INPORT_STRESS_TYPE=1, DT=0.008,
                          Fault_grid=[11, NY_L - 10, 11, NZ_L + 3], mu_s=0.677, mu_d=0.525, Dc=0.4,
                          RS_V0=1e-6, RS_f0=0.6, RS_Vini=1e-12, RS_fw=0.2, RS_L=0.02, smooth_load_T=0.0)
    o = Oracle(p)
    o.set_coord3d(_affine_coords(p, Amap))
    d = {"nx": NX_L + 6, "ny": NY_L + 6, "nz": NZ_L + 6}
    Ai = np.linalg.inv(Amap)
    n = Ai[0] / np.linalg.norm(Ai[0])            # fault normal (grad xi)
    ns = Ai[2] / np.linalg.norm(Ai[2])           # surface normal (grad zeta)
    s1 = Amap[:, 1] / np.linalg.norm(Amap[:, 1])  # along strike (x_eta)
    s2 = np.cross(n, s1)
    z = np.zeros((NZ_L, NY_L))
    T0 = -120e6 * n + 30e6 * s1 + 4e6 * s2       # locked: strength 0.677 * 120 MPa
    stress = {"Tx": z + T0[0], "Ty": z + T0[1], "Tz": z + T0[2], "dTx": z, "dTy": z, "dTz": z}
    if friction == 0:
        stress.update({"mu_s": z + 0.677, "mu_d": z + 0.525, "Dc": z + 0.4, "C0": z})
    else:
        # psi / a = 600: exp(-psi/a) ~ 1e-261 makes the rate-state fault locked
        # to ~1e-240 m/s, and the aging law's healing term exp((f0-psi)/b) ~ 0
        stress.update({"a": z + 0.01, "b": z + 0.012, "L": z + 0.02, "Vw": z + 0.1, "State": z + 6.0})
    for k, v in stress.items():
        o.set_stress(k, v)
    o.setup()
    M = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
    lam, mu, rho = M[10, 10, 10, 10], M[11, 10, 10, 10], M[12, 10, 10, 10]
    X = o.get("C").reshape(3, d["nx"], d["nz"], d["ny"])
    i0 = o.dims()["i0"]
    xc = X[:, i0, d["nz"] - 4, d["ny"] // 2].copy()  # a point of the fault on the surface
    rng = np.random.default_rng(7)
    t1 = np.cross(ns, n) if np.linalg.norm(np.cross(ns, n)) > 1e-6 else s1
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(ns, t1)
    # velocity gradient: traction-free strain rate on the surface plus a spin
    if friction == 0:
        e1, e2, e12 = rng.uniform(-1, 1, 3) * 1e-4
        eps = e1 * np.outer(t1, t1) + e2 * np.outer(t2, t2) + e12 * (np.outer(t1, t2) + np.outer(t2, t1))
        eps += -lam * (e1 + e2) / (lam + 2 * mu) * np.outer(ns, ns)
    else:
        # flat grid: zero stress rate on the surface (z) and on the fault (x)
        eyy = 1e-4
        e = -lam * eyy / (2 * lam + 2 * mu)
        eps = np.diag([e, eyy, e])
    W_ = rng.uniform(-1, 1, 3) * 1e-4
    spin = np.array([[0, -W_[2], W_[1]], [W_[2], 0, -W_[0]], [-W_[1], W_[0], 0]])
    A = eps + spin
    V0 = np.array([0.13, -0.07, 0.05])
    # stress: traction-free on the surface, varying along the surface and in depth
    if friction == 0:
        sig0 = 1.1e6 * np.outer(t1, t1) - 0.6e6 * np.outer(t2, t2) + 0.4e6 * (np.outer(t1, t2) + np.outer(t2, t1))
        u = np.array([0.8, -0.5, 0.3]) * 1e6 / 1e3   # Pa per m of depth
        # depth-linear (zero at the surface), along the surface, and along the
        # fault normal with a nonzero traction on the fault (t2 . n != 0)
        lin = [(ns, np.outer(ns, u) + np.outer(u, ns)), (t1, 0.9e3 * np.outer(t1, t1)),
               (n, 0.6e3 * np.outer(t2, t2))]
    else:
        # normal traction on the fault (x) zero: sxx = sxy = sxz = 0 on x = 0
        sig0 = np.array([[0.0, 0.0, 0.0], [0.0, 1.1e6, 0.0], [0.0, 0.0, 0.0]])
        ex, ey, ez = np.eye(3)
        lin = [(ns, 0.7e3 * (np.outer(ez, ey) + np.outer(ey, ez))    # syz, sxz grow with depth from 0 at the surface
                + 0.3e3 * (np.outer(ex, ez) + np.outer(ez, ex))),
               (ex, 0.5e3 * np.outer(ey, ey)                         # syy and the fault shear sxy along x
                + 0.6e3 * (np.outer(ex, ey) + np.outer(ey, ex))),
               (ey, 0.4e3 * (np.outer(ex, ey) + np.outer(ey, ex)))]  # fault shear sxy along strike
    H_sig = _hooke(0.5 * (A + A.T), lam, mu)
    divs = sum(S @ g for g, S in lin)
    H_V = divs / rho

    def fields(x, t):
        """Exact wavefield at points x [3, ...] and time t: 9 words."""
        dx = x - xc.reshape((3,) + (1,) * (x.ndim - 1))
        V = V0.reshape((3,) + (1,) * (x.ndim - 1)) + np.einsum("ab,b...->a...", A, dx) \
            + t * H_V.reshape((3,) + (1,) * (x.ndim - 1))
        S = np.broadcast_to(sig0 + t * H_sig, x.shape[1:] + (3, 3)).copy()
        for g, T in lin:
            S = S + np.einsum("a...,a->...", dx, g)[..., None, None] * T
        return np.concatenate([V, _voigt(S)])

    return {"o": o, "p": p, "M": M, "X": X, "i0": i0, "d": d, "fields": fields, "H_sig": H_sig, "T0": T0,
            "n": n, "s1": s1, "s2": s2}


def _tractions(Wc, e, J):
    """J (sigma . e) of stress words Wc[3:9] and a metric row e [3, ...]."""
    sxx, syy, szz, sxy, sxz, syz = Wc[3:9]
    return J * np.stack([sxx * e[0] + sxy * e[1] + sxz * e[2],
                         sxy * e[0] + syy * e[1] + syz * e[2],
                         sxz * e[0] + syz * e[1] + szz * e[2]])


def _run_linear(friction, Amap):
    """Set the exact solution at t = 0, advance N_STEPS steps and return the
    largest relative deviations from the exact solution at t = N DT."""
    P = _linear_problem(friction, Amap)
    o, M, X, i0, d, f = P["o"], P["M"], P["X"], P["i0"], P["d"], P["fields"]
    nx, ny, nz = d["nx"], d["ny"], d["nz"]
    nyz = ny * nz
    o.put("W", f(X, 0.0))
    Xf = X[:, i0]                                  # fault plane [3][nz][ny]
    Wf = f(Xf, 0.0)
    e_xi, e_et, e_zt, J = M[0:3, i0], M[3:6, i0], M[6:9, i0], M[9, i0]
    FW = np.zeros((9, 2, nz, ny))
    for m in range(2):
        FW[0:3, m] = Wf[0:3]
        FW[3:6, m] = _tractions(Wf, e_et, J)
        FW[6:9, m] = _tractions(Wf, e_zt, J)
    o.put("FW", FW)
    T1 = np.zeros((3, 7, nz, ny))
    for l in range(7):
        Wl = f(X[:, i0 - 3 + l], 0.0)
        T1[:, l] = _tractions(Wl, M[0:3, i0 - 3 + l], M[9, i0 - 3 + l])
    for a, name in enumerate(("T11", "T12", "T13")):
        o.put(name, T1[a])
    # the traction rate at the fault (read by the DxV_T1 term of slip weakening)
    hT = J * np.einsum("ab,b...->a...", P["H_sig"], e_xi)
    for a, name in enumerate(("hT11", "hT12", "hT13")):
        h = np.zeros((7, nz, ny))
        h[0] = hT[a]
        o.put(name, h)
    for it in range(N_STEPS):
        o.step(it)
    t = N_STEPS * o_dt(P)
    dev = {}
    # volume: planes i0-6..i0+6, the top 10 rows, the clean strike range
    jc = slice(3 + 30 + REACH, 3 + NY_L - 31 - REACH)
    sl = (slice(i0 - 6, i0 + 7), slice(nz - 13, nz - 3), jc)
    Wn = o.get("W").reshape(9, nx, nz, ny)
    We = f(X, t)
    for c in range(9):
        scale = np.abs(We[0:3][(slice(None),) + sl]).max() if c < 3 else np.abs(We[3:9][(slice(None),) + sl]).max()
        dev[f"W{c}"] = np.abs(Wn[c][sl] - We[c][sl]).max() / scale
    # split nodes and fault tractions on the split (not united) nodes of the
    # top 10 rows
    live = np.zeros((nz, ny), bool)
    live[nz - 13:nz - 3, jc] = True
    assert np.all(o.get("united").reshape(NZ_L, NY_L)[live[3:-3, 3:-3]] == 0)
    fs = live
    Wf = f(Xf, t)
    FWn = o.get("FW").reshape(9, 2, nz, ny)
    ex = [Wf[0:3], _tractions(Wf, e_et, J), _tractions(Wf, e_zt, J)]
    for m in range(2):
        for g in range(3):
            for a in range(3):
                c = 3 * g + a
                scale = np.abs(ex[g][a]).max() + 1e-30
                dev[f"FW{c}_{m}"] = np.abs(FWn[c, m][fs] - ex[g][a][fs]).max() / max(scale, np.abs(ex[g]).max())
    T1e = _tractions(Wf, e_xi, J)
    for a, name in enumerate(("T11", "T12", "T13")):
        Tn_ = o.get(name).reshape(7, nz, ny)[3]
        dev[name] = np.abs(Tn_[fs] - T1e[a][fs]).max() / np.abs(T1e).max()
    nj = NY_L
    fs2 = live[3:-3, 3:-3]
    vs = np.hypot(o.get("Vs1"), o.get("Vs2")).reshape(NZ_L, nj)
    dev["slip_rate"] = vs[fs2].max() / np.abs(V_scale(Wf)).max()
    dev["slip"] = np.abs(o.get("slip")).max()
    # RK-weighted stage traction of the last step: the traction at its middle
    Wm = f(Xf, t - 0.5 * o_dt(P))
    n, s1 = P["n"], P["s1"]
    Sm = np.stack([[Wm[3], Wm[6], Wm[7]], [Wm[6], Wm[4], Wm[8]], [Wm[7], Wm[8], Wm[5]]])
    trac = np.einsum("ab...,b->a...", Sm, n) + P["T0"][:, None, None]
    ts1_e = np.einsum("a...,a->...", trac, s1)[3:-3, 3:-3]
    ts1 = o.get("Ts1").reshape(NZ_L, nj)
    dev["Ts1"] = np.abs(ts1[fs2] - ts1_e[fs2]).max() / np.abs(ts1_e[fs2]).max()
    return dev


def o_dt(P):
    return P["p"]["DT"]


def V_scale(Wf):
    return Wf[0:3]


AFFINE = _rotation(0.35, -0.25, 0.3) @ np.diag([1.0, 1.1, 0.9]) + np.array([[0, 0, 0.15], [0, 0, 0], [0, 0, 0]])
FLAT = np.eye(3)

# deviation bounds: FP64 roundoff of a few steps of the scheme
LINEAR_TOL = 1e-9


def _linear_ok(dev):
    bad = {k: v for k, v in dev.items() if not (v <= LINEAR_TOL)}
    return not bad, bad


def test_locked_sw_fault_keeps_linear_solution_exact():
    """Slip weakening, tilted affine grid, free surface: after two steps the
    volume, the split nodes, the fault-plane traction T1 and the RK-weighted
    stage traction equal the exact linear solution; the fault has not
    slipped."""
    ok, bad = _linear_ok(_run_linear(0, AFFINE))
    assert ok, bad


def test_locked_rs_fault_keeps_linear_solution_exact():
    """Rate and state (aging law) in its locked limit on a flat grid with the
    free surface: the trial's traction along the trial direction, the
    effective-normal-stress handling and the writeback reproduce the exact
    linear solution."""
    ok, bad = _linear_ok(_run_linear(1, FLAT))
    assert ok, bad


# planted mistakes (oracle/cgfdm_oracle.cpp MUT(n)) and the pin that must catch them
LINEAR_MUTATIONS_SW = {
    1: "trial formula: sign of the minus-side term",
    2: "trial assembly: plane index of the a_2 term",
    3: "fault_dvelo: constant instead of linear traction extrapolation",
    4: "fault_dvelo: plus-side fault coefficient on the minus side",
    5: "fault_dstrs_f: sign of the 2-4 operator at i0+2",
    6: "fault_dstrs: D matrix of the z derivative on the y derivative",
    7: "fault2wave: metric product in the wrong order",
    8: "fault_rk_stage2: wavefield weight in the stage-traction average",
    10: "wave_deriv: even instead of odd free-surface image",
    11: "wave_deriv: free-surface maps of x and y swapped",
    12: "fault_dstrs_f: sign of the minus-side y term at the surface split node",
    19: "trial writeback: component index of T12",
}


@pytest.mark.parametrize("m", sorted(LINEAR_MUTATIONS_SW))
def test_linear_pin_fails_on_planted_mistake_sw(m):
    set_mutation(m)
    ok, _ = _linear_ok(_run_linear(0, AFFINE))
    assert not ok, LINEAR_MUTATIONS_SW[m]


def test_linear_pin_fails_on_planted_mistake_rs():
    """MUT(9): the rate-state trial takes the s2 component for the s1 one."""
    set_mutation(9)
    ok, _ = _linear_ok(_run_linear(1, FLAT))
    assert not ok


# ---------------------------------------------------------------------------
# State laws at constant slip rate (rk4.cu:791-938)

def _state_oracle(friction, Vslip, psi0, L=0.02, a=0.01, b=0.012, Vw=0.1, DT=0.004):
    NY, NZ = 70, 40
    p = cases.base_params(NX=12, NY=NY, NZ=NZ, usePML=0, FreeSurface=0, NormalRefine=0, Rupture=1,
                          Friction_type=friction, INPORT_STRESS_TYPE=1, DT=DT, Fault_grid=[31, NY - 30, 31, NZ],
                          RS_V0=1e-6, RS_f0=0.6, RS_Vini=1e-12, RS_fw=0.2, RS_L=L)
    o = Oracle(p)
    z = np.zeros((NZ, NY))
    for k, v in {"Tx": z - 1e8, "Ty": z + 3e7, "Tz": z, "dTx": z, "dTy": z, "dTz": z, "a": z + a, "b": z + b,
                 "L": z + L, "Vw": z + Vw, "State": z + psi0}.items():
        o.set_stress(k, v)
    o.setup()
    d = o.dims()
    ny, nz = d["ny"], d["nz"]
    # split-node velocities with a constant slip rate along strike s1
    s1 = o.get("vec_s1").reshape(nz, ny, 3)
    FW = np.zeros((9, 2, nz, ny))
    for c in range(3):
        FW[c, 1] = 0.5 * Vslip * s1[..., c]
        FW[c, 0] = -0.5 * Vslip * s1[..., c]
    o.put("FW", FW)
    live = o.get("united").reshape(NZ, NY) == 0
    return o, live


def _advance_state(o, nsteps):
    for _ in range(nsteps):
        for irk in range(4):
            o.state_rk(irk)


def _rk4_factor(z):
    return 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24


def _aging_dev(mut=0):
    V, psi0, L, b, f0, V0, DT, n = 0.1, 0.75, 0.02, 0.012, 0.6, 1e-6, 0.004, 150
    set_mutation(mut)
    o, live = _state_oracle(1, V, psi0, L=L, b=b, DT=DT)
    _advance_state(o, n)
    t = n * DT
    theta0 = L / V0 * math.exp((psi0 - f0) / b)
    theta = L / V + (theta0 - L / V) * math.exp(-V * t / L)
    psi = f0 + b * math.log(V0 * theta / L)
    st = o.get("State").reshape(live.shape)[live]
    slip = o.get("slip").reshape(live.shape)[live]
    set_mutation(0)
    return max(np.abs(st - psi).max() / abs(psi), np.abs(slip - V * t).max() / (V * t))


def test_aging_law_constant_rate_closed_form():
    """Aging law at constant V: theta = L/V + (theta0 - L/V) e^(-Vt/L) with
    psi = f0 + b ln(V0 theta / L); RK4 with V DT / L = 0.02 stays within 1e-9,
    and slip = V t exactly."""
    assert _aging_dev() < 1e-9


def test_aging_pin_fails_on_planted_mistake():
    """MUT(13): the sign of the slip-rate term."""
    assert _aging_dev(13) > 1e-3


def _slip_law_fss(V, mut=0):
    """Steady-state friction implied by the slip law's fixed point: RK4 of
    psi' = -V/L (psi - psi_ss) is psi_n = psi_ss + (psi0 - psi_ss) R(z)^n,
    z = -V DT/L, so two runs give psi_ss; the rate-state friction law then
    gives f_ss = a asinh(V/(2 V0) e^(psi_ss/a))."""
    a, L, DT, V0, n = 0.01, 0.02, 0.004, 1e-6, 20
    set_mutation(mut)
    vals = []
    for psi0 in (0.5, 0.9):
        o, live = _state_oracle(2, V, psi0, L=L, a=a, DT=DT)
        _advance_state(o, n)
        vals.append((psi0, o.get("State").reshape(live.shape)[live].mean()))
    set_mutation(0)
    Rn = _rk4_factor(-V * DT / L) ** n
    (p0, pa), (p1, pb) = vals
    # psi_n = psi_ss (1 - Rn) + psi0 Rn for both starts: consistent fixed point
    ss_a = (pa - p0 * Rn) / (1 - Rn)
    ss_b = (pb - p1 * Rn) / (1 - Rn)
    return a * math.asinh(V / (2 * V0) * math.exp(ss_a / a)), ss_a, ss_b


def test_slip_law_steady_state_limits():
    """Slip law with flash weakening (rk4.cu:902-911): the fixed point of the
    state equation is the steady state whose friction is
    f0 - (b - a) ln(V/V0) for V << Vw and falls towards fw for V >> Vw
    (as fw + (f_lv - fw) Vw/V)."""
    a, b, f0, V0, fw, Vw = 0.01, 0.012, 0.6, 1e-6, 0.2, 0.1
    V = 1e-4 * Vw
    fss, sa, sb = _slip_law_fss(V)
    assert abs(sa - sb) < 1e-9
    assert abs(fss - (f0 - (b - a) * math.log(V / V0))) < 1e-9
    V = 100 * Vw
    fss, sa, sb = _slip_law_fss(V)
    flv = f0 - (b - a) * math.log(V / V0)
    assert abs(fss - (fw + (flv - fw) * Vw / V)) < 1e-6 * fss


def test_slip_law_pin_fails_on_planted_mistake():
    """MUT(14): a and b swapped in the low-velocity steady state."""
    a, b, f0, V0, Vw = 0.01, 0.012, 0.6, 1e-6, 0.1
    V = 1e-4 * Vw
    fss, _, _ = _slip_law_fss(V, mut=14)
    assert abs(fss - (f0 - (b - a) * math.log(V / V0))) > 1e-4


# ---------------------------------------------------------------------------
# ADE-PML auxiliary ODE (pmlx.cu:126-210)

def _pml_problem(free=0, bmax=3.0, Amap=None, seed=0, linear=True):
    p = cases.base_params(NX=30, NY=26, NZ=24, usePML=1, PML_N=8, FreeSurface=free, NormalRefine=0,
                          PML_bmax=bmax, Rupture=0, DT=0.004)
    o = Oracle(p)
    if Amap is not None:
        o.set_coord3d(_affine_coords(p, Amap))
    o.setup()
    d = o.dims()
    nx, ny, nz = d["nx"], d["ny"], d["nz"]
    X = o.get("C").reshape(3, nx, nz, ny)
    W = np.zeros((9, nx, nz, ny))
    if linear:
        g = np.array([0.3, -0.2, 0.1, 2e6, -1e6, 0.5e6, 0.7e6, -0.3e6, 0.2e6]) / 1e3
        for c in range(9):
            W[c] = g[c] * X[0]
    else:
        W = cases.random_wavefield(nx, ny, nz, seed=seed)
    return o, W, d


def _pml_dev(mut=0):
    set_mutation(mut)
    o, W, d = _pml_problem()
    nx, ny, nz, ni, nj, nk = d["nx"], d["ny"], d["nz"], d["ni"], d["nj"], d["nk"]
    N, DT, nsteps = 8, 0.004, 12
    M = o.get("M").reshape(13, nx, nz, ny)
    lam, mu, rho = M[10, 10, 10, 10], M[11, 10, 10, 10], M[12, 10, 10, 10]
    for it in range(nsteps):
        for irk in range(4):
            flag = 1 if irk % 2 == 0 else -1
            o.put("W", W)
            o.put("hW", np.zeros_like(W))
            o.pml_deriv(0, 0, flag)
            o.pml_deriv(0, 1, flag)
            o.wave_rk(irk)
    aux = o.get("PMLx1").reshape(9, N, nk, nj)
    set_mutation(0)
    # the x-axis source of the linear field (identity metric): the
    # derivatives along x only
    g = (W[:, 20, 10, 10] - W[:, 19, 10, 10]) / o_dh()
    Dr = np.array([g[3] / rho, g[6] / rho, g[7] / rho, (lam + 2 * mu) * g[0], lam * g[0], lam * g[0],
                   mu * g[1], mu * g[2], 0.0])
    A, D = o.get("Ax"), o.get("Dx")
    t = nsteps * DT
    worst = 0.0
    sc = max(np.abs(D[li] * Dr / (A[li] + D[li])).max() for li in range(N))
    for li in range(N):
        zz = -(A[li] + D[li]) * DT
        ystar = D[li] * Dr / (A[li] + D[li])
        y_rk = ystar * (1 - _rk4_factor(zz) ** nsteps)
        y_exact = ystar * (1 - math.exp(-(A[li] + D[li]) * t))
        got = aux[:, li, 8:-8, 8:-8].reshape(9, -1)
        worst = max(worst, np.abs(got - y_rk[:, None]).max() / sc)
        assert np.abs(y_rk - y_exact).max() < 1e-3 * sc  # RK4 of the same ODE
    return worst


def o_dh():
    return 100.0


def test_pml_aux_ode_closed_form():
    """x-face auxiliary fields under a time-constant x derivative: the RK4
    solution y_n = y* (1 - R(z)^n) of y' = -(alpha + d) y + d D,
    y* = d D / (alpha + d), z = -(alpha + d) DT, per layer of the face."""
    assert _pml_dev() < 1e-10


def test_pml_pin_fails_on_planted_mistake():
    """MUT(15): the sign of d in the decay rate."""
    assert _pml_dev(15) > 1e-3


def test_pml_aux_zero_source_stays_zero():
    o, W, d = _pml_problem(linear=True)
    W[:] = 0.0
    for irk in range(4):
        o.put("W", W)
        o.put("hW", np.zeros_like(W))
        o.pml_deriv(0, 0, 1 if irk % 2 == 0 else -1)
        o.wave_rk(irk)
    assert np.abs(o.get("PMLx1")).max() == 0.0


def _pml_surface_traction(mut=0):
    """Top-row source of the x-face auxiliary fields with the free surface
    and beta = 1 (PML_bmax = 1): the surface correction (pmlx.cu:185-210)
    adds the stress rate of the z derivative the traction-free condition
    implies, so the source's traction on the surface vanishes."""
    set_mutation(mut)
    R = _rotation(0.2, -0.15, 0.1) @ np.diag([1.0, 1.05, 0.95])
    o, W, d = _pml_problem(free=1, bmax=1.0, Amap=R, linear=False, seed=3)
    nx, ny, nz, nj, nk = d["nx"], d["ny"], d["nz"], d["nj"], d["nk"]
    o.put("W", W)
    o.put("hW", np.zeros_like(W))
    o.pml_deriv(0, 0, 1)
    o.wave_rk(0)  # aux = 0 + DT/2 * source
    set_mutation(0)
    S = o.get("PMLx1").reshape(9, 8, nk, nj)[:, :, nk - 1, 3:-3] / (0.5 * 0.004)
    M = o.get("M").reshape(13, nx, nz, ny)
    e_z = M[6:9, 10, nz - 4, 10]
    t = np.stack([S[3] * e_z[0] + S[6] * e_z[1] + S[7] * e_z[2],
                  S[6] * e_z[0] + S[4] * e_z[1] + S[8] * e_z[2],
                  S[7] * e_z[0] + S[8] * e_z[1] + S[5] * e_z[2]])
    return np.abs(t).max() / np.abs(S[3:]).max()


def test_pml_surface_correction_is_traction_free():
    assert _pml_surface_traction() < 1e-12


def test_pml_surface_pin_fails_on_planted_mistake():
    """MUT(16): the z components paired wrongly in the shear term (the pin's
    bound is 1e-12)."""
    assert _pml_surface_traction(16) > 1e-8


# ---------------------------------------------------------------------------
# Rupture sensor (rup_sensor.cu:7-97)

def _sensor(mut=0):
    NY, NZ = 70, 44
    case = cases.sw_case(NX=12, NY=NY, NZ=NZ)
    o = Oracle(case["params"])
    cases.apply_to_oracle(case, o)
    o.setup()
    cy, cz, cyz = 2.0e4, -1.2e4, 0.7e4
    j = np.arange(NY)[None, :]
    k = np.arange(NZ)[:, None]
    ts1 = 6e7 + 3e5 * j - 2e5 * k + cy * j * j + cz * k * k + cyz * j * k
    o.put("Ts1", ts1.ravel())
    set_mutation(mut)
    o.rup_sensor()
    set_mutation(0)
    rs = o.get("rup_sensor").reshape(NZ, NY)
    fg = case["params"]["Fault_grid"]
    # nodes whose neighbours' second differences are computed too
    jj = slice(fg[0] + 3, fg[1] - 4)
    kk = slice(max(fg[2] + 3, 31), fg[3] - 4)
    # second differences: -0.25 * 2 cy and -0.25 * 2 cz; the sensor pairs
    # the z difference at the node with the y differences above and below
    mg = 0.25 * (cy - cz) ** 2
    expect = mg / (ts1 ** 2 + 1e-16) + 1e-16
    return np.abs(rs[kk, jj] - expect[kk, jj]).max() / np.abs(expect[kk, jj]).max()


def test_rup_sensor_quadratic_closed_form():
    assert _sensor() < 1e-12


def test_rup_sensor_pin_fails_on_planted_mistake():
    """MUT(17): the z difference paired with z differences."""
    assert _sensor(17) > 0.5


# ---------------------------------------------------------------------------
# Free-surface split-node maps (fault_coef.cu:855-1010)

def _surface_maps(mut=0):
    NX, NY, NZ = 26, 70, 40
    case = cases.sw_case(NX=NX, NY=NY, NZ=NZ, curved=True)
    # a material contrast across the fault (Media3D, media.cu:140-204)
    minus = (np.arange(NX) < NX // 2)[:, None, None] * np.ones((1, NZ, NY), bool)
    vp, vs, rho = (np.where(minus, a, b) for a, b in ((6000.0, 5200.0), (3464.0, 3000.0), (2670.0, 2500.0)))
    set_mutation(mut)
    o = Oracle(case["params"])
    cases.apply_to_oracle(case, o)
    o.set_media3d(vp, vs, rho)
    o.setup()
    set_mutation(0)
    d = o.dims()
    ny, nz = d["ny"], d["nz"]
    k = nz - 4
    g = lambda n: o.get(n).reshape(-1, 9)  # noqa: E731
    P = [g(f"matPlus2Min{t}f").reshape(ny, 3, 3) for t in (1, 2, 3)]
    Q = [g(f"matMin2Plus{t}f").reshape(ny, 3, 3) for t in (1, 2, 3)]
    VX = [o.get(f"matVx2Vz{m}").reshape(ny, 3, 3) for m in (1, 2)]
    VY = [o.get(f"matVy2Vz{m}").reshape(ny, 3, 3) for m in (1, 2)]
    D = {(m, a, b): o.get(f"D{a}{b}_{m}").reshape(nz, ny, 3, 3) for m in (1, 2) for a in (1, 2, 3) for b in (1, 2, 3)}
    rng = np.random.default_rng(5)
    round_trip, free = 0.0, 0.0
    for j in range(8, ny - 8, 5):
        dxp, dyp, dym = rng.standard_normal((3, 3))
        dxm = P[0][j] @ dxp + P[1][j] @ dyp - P[2][j] @ dym
        back = Q[0][j] @ dxm + Q[1][j] @ dym - Q[2][j] @ dyp
        round_trip = max(round_trip, np.abs(back - dxp).max() / np.abs(dxp).max())
        for m, (dx, dy) in enumerate(((dxm, dym), (dxp, dyp))):
            dz = VX[m][j] @ dx + VY[m][j] @ dy
            tz = D[(m + 1, 3, 1)][k, j] @ dx + D[(m + 1, 3, 2)][k, j] @ dy + D[(m + 1, 3, 3)][k, j] @ dz
            sc = np.abs(D[(m + 1, 3, 3)][k, j]).max() * np.abs(dz).max()
            free = max(free, np.abs(tz).max() / sc)
    return round_trip, free


def test_surface_split_node_maps():
    """Bi-material curved grid: on each side the map to DzV makes the
    zeta traction rate vanish, and the minus-to-plus map undoes the
    plus-to-minus map."""
    rt, fr = _surface_maps()
    assert rt < 1e-10 and fr < 1e-10, (rt, fr)


def test_surface_maps_pin_fails_on_planted_mistake():
    """MUT(18): the minus-to-plus x map multiplied in the wrong order."""
    rt, _ = _surface_maps(18)
    assert rt > 1e-3
