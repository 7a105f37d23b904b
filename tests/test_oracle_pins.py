"""Pins of the FP64 oracle against mathematics (CPU).

The reference ships no golden vectors or tests for this path (its
validation data is not in the repository), so each oracle function is
pinned by a closed form, an invariant or a brute-force computation chosen so
that a dropped term, a wrong sign/index or a transposed operand fails.
"""
import math

import numpy as np
import pytest

import cases
from oracle.binding import Oracle

# DRP coefficients as printed in the reference (macdrp.h:23-27)
C = np.array([0.29164704779069994, 0.66674514216519998, 1.2501177132452999,
              0.33341180882999999, 0.041686285540599999])


def _oracle(params, geom=None, coord3d=None):
    o = Oracle(params)
    if geom is not None:
        o.set_geometry(*geom)
    if coord3d is not None:
        o.set_coord3d(coord3d)
    o.setup()
    return o


def _flat(**kw):
    p = dict(usePML=0, FreeSurface=0, NormalRefine=0)
    p.update(kw)
    return cases.base_params(NX=20, NY=20, NZ=20, **p)


def _grid_xyz(o):
    d = o.dims()
    C3 = o.get("C").reshape(3, d["nx"], d["nz"], d["ny"])
    return d, C3


def test_drp_coefficient_moments():
    """Forward operator -c1 f(-1) - c2 f0 + c3 f1 - c4 f2 + c5 f3: zeroth
    moment 0, first moment 1 (consistency of a first derivative)."""
    w = np.array([-C[0], -C[1], C[2], -C[3], C[4]])
    x = np.arange(-1, 4)
    assert abs(w.sum()) < 1e-15
    assert abs((w * x).sum() - 1.0) < 1e-11  # printed coefficients: 1 - 2.2e-12


def test_wave_deriv_linear_field_exact():
    """Identity metric, V = (a x, b y, c z) + linear stresses: the interior
    derivatives are exact, so hT = Hooke(grad V) and hV = div(sigma)/rho."""
    o = _oracle(_flat())
    d, X = _grid_xyz(o)
    M = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
    s = (slice(3, -3),) * 3
    assert np.allclose(M[0][s], 1.0) and np.allclose(M[4][s], 1.0) and np.allclose(M[8][s], 1.0)
    assert np.allclose(M[1][s], 0.0) and np.allclose(M[9][s], 1.0)
    x, y, z = X[0], X[1], X[2]
    a, b, c = 0.3, -0.7, 0.2
    W = np.zeros((9,) + x.shape)
    W[0], W[1], W[2] = a * x, b * y, c * z
    W[3] = 5e3 * x  # Txx grows along x
    W[6] = -2e3 * y  # Txy along y
    W[8] = 7e3 * z  # Tyz along z
    W[:, :3, :, :] = W[:, -3:, :, :] = 0
    W[:, :, :3, :] = W[:, :, -3:, :] = 0
    W[:, :, :, :3] = W[:, :, :, -3:] = 0
    lam, mu, rho = M[10][10, 10, 10], M[11][10, 10, 10], M[12][10, 10, 10]
    for flags in [(1, 1, 1), (-1, -1, -1), (1, -1, 1)]:
        o.put("W", W)
        o.wave_deriv(*flags)
        h = o.get("hW").reshape(9, d["nx"], d["nz"], d["ny"])
        core = (slice(9, -9),) * 3  # stencils stay off the zero ghosts
        np.testing.assert_allclose(h[3][core], (lam + 2 * mu) * a + lam * (b + c), rtol=1e-10)
        np.testing.assert_allclose(h[4][core], (lam + 2 * mu) * b + lam * (a + c), rtol=1e-10)
        np.testing.assert_allclose(h[5][core], (lam + 2 * mu) * c + lam * (a + b), rtol=1e-10)
        np.testing.assert_allclose(h[6][core], 0.0, atol=1e-12 * (lam + 2 * mu))
        np.testing.assert_allclose(h[0][core], (5e3 - 2e3) / rho, rtol=1e-10)  # dTxx/dx + dTxy/dy
        np.testing.assert_allclose(h[1][core], 7e3 / rho, rtol=1e-10)  # dTyz/dz
        np.testing.assert_allclose(h[2][core], 0.0, atol=1e-9)  # dTyz/dy = 0


def test_drp_quadratic_exact_and_backward_is_mirror():
    """The one-sided DRP operator of this reference has zero second moment
    (exact on quadratics); the backward operator LB is the mirror image of
    LF: B(f)(i) = -F(f reversed)(n-1-i), checked on a random field so a
    wrong or transposed backward coefficient fails."""
    w = np.array([-C[0], -C[1], C[2], -C[3], C[4]])
    assert abs((w * np.arange(-1, 4) ** 2).sum()) < 1e-11
    assert abs((w * np.arange(-1, 4) ** 3).sum()) < 1e-10  # 4th-order one-sided scheme
    assert abs((w * np.arange(-1, 4) ** 4).sum()) > 1e-2  # not exact on quartics
    o = _oracle(_flat())
    d, X = _grid_xyz(o)
    rng = np.random.default_rng(3)
    W = np.zeros((9, d["nx"], d["nz"], d["ny"]))
    W[0, 3:-3, 3:-3, 3:-3] = rng.standard_normal((d["ni"], d["nk"], d["nj"]))
    o.put("W", W)
    o.wave_deriv(-1, 1, 1)
    hb = o.get("hW").reshape(9, d["nx"], d["nz"], d["ny"])[3]
    o.put("W", W[:, ::-1].copy())
    o.wave_deriv(1, 1, 1)
    hf = o.get("hW").reshape(9, d["nx"], d["nz"], d["ny"])[3][::-1]
    np.testing.assert_allclose(hb[3:-3], -hf[3:-3], rtol=1e-12, atol=1e-12 * np.abs(hb).max())
    # and F is exact on a quadratic away from the ghosts
    o.put("W", np.where(np.arange(9)[:, None, None, None] == 0, (X[0] / 100.0) ** 2, 0.0) * (np.abs(W[:1]) >= 0))
    o.wave_deriv(1, 1, 1)
    h = o.get("hW").reshape(9, d["nx"], d["nz"], d["ny"])
    M = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
    l2m = M[10][10, 10, 10] + 2 * M[11][10, 10, 10]
    core = (slice(9, -9),) * 3
    np.testing.assert_allclose(h[3][core], l2m * 2 * X[0][core] / 100.0 ** 2, rtol=1e-9, atol=1e-9 * l2m)


def test_metric_normal_refine_closed_form():
    """NormalRefine spacing (coord.cu:175-209): x_xi = (x(i+1)-x(i-1))/(2 DH),
    xi_x = 1/x_xi, J = x_xi, with the cosine taper built independently."""
    p = _flat(NormalRefine=1)
    p["NX"] = 140
    o = _oracle(p)
    d, X = _grid_xyz(o)
    NX, DH = p["NX"], p["DH"]
    ic = NX // 2 + 3
    xl = np.zeros(NX + 6)
    for i in range(ic + 1, NX + 6):
        dist = i - ic
        cp = 0.0 if dist < 10 else (1 - math.cos(math.pi * (i - ic - 10) / 45.0) if dist < 55 else 2.0)
        xl[i] = xl[i - 1] + DH * (0.5 + 0.25 * cp)
    for i in range(ic - 1, -1, -1):
        dist = ic - i
        cp = 0.0 if dist < 10 else (1 - math.cos(math.pi * (i - ic + 10) / 45.0) if dist < 55 else 2.0)
        xl[i] = xl[i + 1] - DH * (0.5 + 0.25 * cp)
    np.testing.assert_allclose(X[0][3:-3, 10, 10], xl[3:-3], rtol=0, atol=1e-9)
    M = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
    xxi = (xl[4:-2] - xl[2:-4]) / (2 * DH)
    np.testing.assert_allclose(M[9][3:-3, 10, 10], xxi, rtol=1e-12)
    np.testing.assert_allclose(M[0][3:-3, 10, 10], 1 / xxi, rtol=1e-12)
    # spacing: 0.5 DH at the fault, DH far away
    assert abs((xl[ic + 1] - xl[ic]) - 0.5 * DH) < 1e-9
    assert abs((xl[ic + 60] - xl[ic + 59]) - DH) < 1e-9


def _rotation(a, b, c):
    ca, sa, cb, sb, cc, sc = math.cos(a), math.sin(a), math.cos(b), math.sin(b), math.cos(c), math.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def _rotated_coords(p, R, stretch=(1.0, 1.0, 1.0)):
    nx, ny, nz = p["NX"] + 6, p["NY"] + 6, p["NZ"] + 6
    I, K, J = np.meshgrid(np.arange(nx), np.arange(nz), np.arange(ny), indexing="ij")
    idx = np.stack([I * stretch[0], J * stretch[1], K * stretch[2]]) * p["DH"]
    xyz = np.einsum("ab,bijk->aijk", R, idx)
    return xyz


def test_metric_rotated_affine_grid():
    """x = R (i,j,k) DH: grad xi = rows of R^-1 = columns of R, J = det R."""
    p = _flat()
    R = _rotation(0.4, -0.3, 0.9)
    o = _oracle(p, coord3d=_rotated_coords(p, R))
    d = o.dims()
    M = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
    s = (slice(3, -3),) * 3
    Ri = np.linalg.inv(R)
    for a in range(3):
        for b in range(3):
            np.testing.assert_allclose(M[3 * a + b][s], Ri[a, b], atol=1e-12)
    np.testing.assert_allclose(M[9][s], np.linalg.det(R), rtol=1e-12)


def test_wave_deriv_rotation_invariance():
    """The same physical linear velocity field on a rotated, stretched grid
    gives the same stress rates (chain rule through the metric)."""
    p = _flat()
    G = np.array([[0.3, -0.1, 0.05], [0.2, -0.4, 0.1], [-0.15, 0.25, 0.6]])  # dV_p/dx_q
    lam_mu = None
    res = []
    for R, st in [(np.eye(3), (1.0, 1.0, 1.0)), (_rotation(0.5, 0.2, -0.7), (1.0, 1.3, 0.8))]:
        xyz = _rotated_coords(p, R, st)
        o = _oracle(p, coord3d=xyz)
        d = o.dims()
        M = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
        lam_mu = (M[10][10, 10, 10], M[11][10, 10, 10])
        W = np.zeros((9, d["nx"], d["nz"], d["ny"]))
        for q in range(3):
            W[q] = np.einsum("b,bijk->ijk", G[q], xyz)
        W[:, :3] = W[:, -3:] = 0
        W[:, :, :3] = W[:, :, -3:] = 0
        W[:, :, :, :3] = W[:, :, :, -3:] = 0
        o.put("W", W)
        o.wave_deriv(1, -1, 1)
        h = o.get("hW").reshape(9, d["nx"], d["nz"], d["ny"])
        res.append(h[3:, 10, 10, 10])
    lam, mu = lam_mu
    eps = 0.5 * (G + G.T)
    sig = lam * np.trace(G) * np.eye(3) + 2 * mu * eps
    expect = np.array([sig[0, 0], sig[1, 1], sig[2, 2], sig[0, 1], sig[0, 2], sig[1, 2]])
    np.testing.assert_allclose(res[0], expect, rtol=1e-10)
    np.testing.assert_allclose(res[1], expect, rtol=1e-9)


def test_pml_profiles_closed_form():
    """abs_init (init_pml.cu:18-155): d = dmax r^2 / beta, a = amax (1-r),
    beta = 1 + (bmax-1) r^2 with r = distance/L0, L0 = DH (N-1)."""
    p = cases.base_params(NX=40, NY=30, NZ=30, PML_N=10, PML_velocity=5000.0, PML_bmax=2.5, PML_fc=1.2)
    o = _oracle(p)
    N, DH = 10, p["DH"]
    L0 = DH * (N - 1)
    R = 10 ** (-((math.log10(N) - 1) / math.log10(2) + 3))
    dmax = -5000.0 / (2 * L0) * math.log(R) * 3
    amax = math.pi * 1.2
    Bx, Ax, Dx = o.get("Bx"), o.get("Ax"), o.get("Dx")
    for i in range(40):
        if i < N:
            r = (N - 1 - i) * DH / L0
        elif i >= 40 - N:
            r = (i - (40 - N)) * DH / L0
        else:
            assert Bx[i] == 1 and Ax[i] == 0 and Dx[i] == 0
            continue
        beta = 1 + 1.5 * r * r
        assert abs(Bx[i] - beta) < 1e-12
        assert abs(Ax[i] - amax * (1 - r)) < 1e-12
        assert abs(Dx[i] - dmax * r * r / beta) < 1e-9
    # free surface: no PML on the top face
    Bz = o.get("Bz")
    assert np.all(Bz[-N:] == 1.0) and Bz[0] > 1.0


def test_cerjan_profile_min_of_faces():
    """init_cerjan_host (cerjan.cu:31-164): exp(-(alpha (ND-d))^2) per face,
    alpha = sqrt(-ln 0.92)/ND, combined by min; no damping on a free top."""
    p = cases.base_params(NX=30, NY=26, NZ=24, usePML=0, useCerjan=1, DAMP_N=6)
    o = _oracle(p)
    d = o.dims()
    damp = o.get("damp").reshape(d["nx"], d["nz"], d["ny"])
    ND = 6
    al = math.sqrt(-math.log(0.92)) / ND

    def prof(n, lo, hi):
        v = np.ones(n + 6)
        for q in range(ND):
            if lo:
                v[3 + q] = math.exp(-(al * (ND - q)) ** 2)
            if hi:
                v[n + 2 - q] = math.exp(-(al * (ND - q)) ** 2)
        return v
    px, py, pz = prof(30, 1, 1), prof(26, 1, 1), prof(24, 1, 0)
    expect = np.minimum(px[:, None, None], np.minimum(pz[None, :, None], py[None, None, :]))
    s = (slice(3, -3),) * 3
    np.testing.assert_allclose(damp[s], expect[s], rtol=1e-12)


def test_rk4_linear_ode():
    """RK_Syn (rk4.cu:20-33) on W' = lam W reproduces the RK4 stability
    polynomial 1 + z + z^2/2 + z^3/6 + z^4/24 exactly."""
    p = _flat()
    o = _oracle(p)
    d = o.dims()
    n = 9 * d["nx"] * d["ny"] * d["nz"]
    rng = np.random.default_rng(0)
    W0 = rng.standard_normal(n)
    lam = -3.7
    o.put("W", W0)
    for irk in range(4):
        o.put("hW", lam * o.get("W"))
        o.wave_rk(irk)
    z = lam * p["DT"]
    amp = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    np.testing.assert_allclose(o.get("W"), amp * W0, rtol=1e-13)


@pytest.mark.parametrize("psi,a,vhat,coef", [(0.75, 0.01, 1e-9, 1e-3), (0.6, 0.016, 2.5, 0.8),
                                             (0.82, 0.008, 1e-3, 4e-2), (0.7, 0.012, 10.0, 3.0)])
def test_newton_root_vs_bisection(psi, a, vhat, coef):
    """NRsearch (trial_ratestate.cu:20-57) converges to the root of
    2 V0 exp(-psi/a) sinh x + coef x = Vhat found by bisection."""
    o = _oracle(_flat())
    V0 = 1e-6
    f = lambda x: 2 * V0 * math.exp(-psi / a) * math.sinh(x) + coef * x - vhat
    lo, hi = 0.0, 1.0
    while f(hi) < 0:
        hi *= 2
    for _ in range(300):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    x_star = 0.5 * (lo + hi)
    x0 = math.asinh(math.exp(psi / a) * vhat / (2 * V0))
    x, it = o.nr_search(x0, psi, a, vhat, coef)
    assert it <= 100
    assert abs(x - x_star) <= 1e-9 * max(1.0, abs(x_star))


def test_free_surface_maps_are_traction_free():
    """matVx2Vz/matVy2Vz (fault_coef.cu:1320-1486): with DzV = A DxV + B DyV
    the traction rate on the zeta surface, computed from the strain rate
    with numpy, vanishes."""
    p = cases.base_params(NX=24, NY=30, NZ=24, usePML=0, INPORT_GRID_TYPE=1)
    geom = cases.bent_geometry(30, 24, 100.0, amp=0.4, seed=4)
    o = _oracle(p, geom=geom)
    d = o.dims()
    M = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
    A = o.get("matVx2Vz").reshape(d["nx"], d["ny"], 3, 3)
    B = o.get("matVy2Vz").reshape(d["nx"], d["ny"], 3, 3)
    rng = np.random.default_rng(1)
    k = d["nz"] - 4
    for (i, j) in [(5, 7), (12, 15), (20, 25)]:
        e = M[:9, i, k, j].reshape(3, 3)  # rows: grad xi, grad eta, grad zeta
        lam, mu = M[10, i, k, j], M[11, i, k, j]
        dxv, dyv = rng.standard_normal(3), rng.standard_normal(3)
        dzv = A[i, j] @ dxv + B[i, j] @ dyv
        G = np.outer(dxv, e[0]) + np.outer(dyv, e[1]) + np.outer(dzv, e[2])  # dV_p/dx_q
        sig = lam * np.trace(G) * np.eye(3) + mu * (G + G.T)
        t = sig @ e[2]
        assert np.linalg.norm(t) < 1e-9 * (lam + 2 * mu) * np.linalg.norm(G)


def test_fault_D_matrices_are_traction_rates():
    """D_ab (fault_coef.cu:236-426): J grad_xi_a . sigma_dot equals
    sum_b D_ab dV/dxi_b for a random velocity gradient (numpy Hooke's law),
    on a curved grid; and the DxV_T1 maps are D11^-1, D11^-1 D12, D11^-1 D13."""
    p = cases.sw_case(curved=True)["params"]
    case = cases.sw_case(curved=True)
    o = Oracle(p)
    cases.apply_to_oracle(case, o)
    o.setup()
    d = o.dims()
    nyz = d["ny"] * d["nz"]
    M = o.get("M").reshape(13, d["nx"], d["nz"], d["ny"])
    Dm = {}
    for m in (1, 2):
        for a in range(3):
            for b in range(3):
                Dm[(m, a, b)] = o.get(f"D{a+1}{b+1}_{m}").reshape(d["nz"], d["ny"], 3, 3)
    lam_f = o.get("rho_f")  # noqa: F841  (rho_f fetched to check the field exists)
    rng = np.random.default_rng(2)
    i0 = d["i0"]
    for (k, j) in [(35, 40), (44, 33), (20, 60)]:
        e = M[:9, i0, k, j].reshape(3, 3)
        J = M[9, i0, k, j]
        for m, side in ((1, -1), (2, 1)):
            lam, mu = M[10, i0 + side, k, j], M[11, i0 + side, k, j]
            dV = rng.standard_normal((3, 3))  # dV[b] = dV/dxi_b
            G = sum(np.outer(dV[b], e[b]) for b in range(3))
            sig = lam * np.trace(G) * np.eye(3) + mu * (G + G.T)
            for a in range(3):
                lhs = J * sig @ e[a]
                rhs = sum(Dm[(m, a, b)][k, j] @ dV[b] for b in range(3))
                np.testing.assert_allclose(rhs, lhs, rtol=1e-10, atol=1e-6 * np.abs(lhs).max())
        T1m = o.get("matT1toVxm").reshape(d["nz"], d["ny"], 3, 3)[k, j]
        Vym = o.get("matVytoVxm").reshape(d["nz"], d["ny"], 3, 3)[k, j]
        D11 = Dm[(1, 0, 0)][k, j]
        np.testing.assert_allclose(T1m @ D11, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(D11 @ Vym, Dm[(1, 0, 1)][k, j], rtol=1e-9, atol=1e-9 * np.abs(D11).max())
    assert nyz > 0


def test_coulomb_cap_holds_on_sliding_nodes():
    """trial_sw (trial_slipweakening.cu:513-615): after rupture, the trial
    traction of sliding nodes sits exactly on tau_c = -mu(D) tau_n + C0 and
    no node exceeds it."""
    case = cases.sw_case(NX=26, NY=72, NZ=44)
    o = Oracle(case["params"])
    cases.apply_to_oracle(case, o)
    o.setup()
    for it in range(25):
        o.step(it)
    ts = np.hypot(o.get("tTs1"), o.get("tTs2"))
    tn = o.get("tTn")
    slip = o.get("slip")
    flag = o.get("flag_rup")
    united = o.get("united")
    mus = o.get("str_peak")
    Dc = case["stress"]["Dc"].ravel()
    mu_d = case["params"]["mu_d"]
    fr = np.where(slip <= Dc, mus - (mus - mu_d) * slip / Dc, mu_d)
    tc = -fr * tn
    live = united == 0
    assert np.any(flag[live] == 1)
    slide = live & (flag == 1) & (tn < 0)
    np.testing.assert_allclose(ts[slide], tc[slide], rtol=1e-9)
    stick = live & (flag == 0)
    assert np.all(ts[stick] <= tc[stick] * (1 + 1e-12))


def test_locked_fault_stays_locked():
    """Prestress below strength everywhere, no nucleation, zero initial
    wavefield: nothing moves (split-node and volume fields stay zero)."""
    case = cases.sw_case(NX=24, NY=70, NZ=40)
    ts = case["stress"]["Ty"]
    case["stress"]["Ty"] = np.maximum(ts, -60e6)
    o = Oracle(case["params"])
    cases.apply_to_oracle(case, o)
    o.setup()
    for it in range(6):
        o.step(it)
    assert np.max(np.abs(o.get("slip"))) == 0.0
    assert np.max(np.abs(o.get("FW"))) == 0.0
    assert np.max(np.abs(o.get("W"))) == 0.0


def _smooth_oracle(R):
    NY, NZ = 40, 36
    p = cases.base_params(NX=20, NY=NY, NZ=NZ, usePML=0, FreeSurface=0, NormalRefine=0, Rupture=1,
                          INPORT_STRESS_TYPE=1, FaultSmooth=1, Fault_grid=[4, NY - 3, 4, NZ - 3])
    o = Oracle(p)
    o.set_coord3d(_rotated_coords(p, R))
    z = np.zeros((NZ, NY))
    for k, v in {"Tx": -1e8 + z, "Ty": 5e7 + z, "Tz": z, "dTx": z, "dTy": z, "dTz": z,
                 "mu_s": 0.6 + z, "mu_d": 0.5 + z, "Dc": 0.4 + z, "C0": z}.items():
        o.set_stress(k, v)
    o.setup()
    return o


def test_smooth_uniform_stress_is_fixed_point():
    """smooth_gauss_volume1 (smooth.cu:64-296) on a rotated affine grid: with
    a uniform stress in the volume and split-node tractions equal to its
    conservative tractions J sigma.grad(eta), J sigma.grad(zeta), the
    normalised half-volume average returns the same tractions on both sides
    (weights of each side sum to 1; traction = S . (grad eta) column, which a
    transposed product E.S would break), and the slip-weighted blend leaves
    them unchanged."""
    R = _rotation(0.35, -0.2, 0.5)
    o = _smooth_oracle(R)
    d = o.dims()
    nx, ny, nz = d["nx"], d["ny"], d["nz"]
    sig = np.array([[3.0e7, 1.1e7, -0.7e7], [1.1e7, -2.0e7, 0.4e7], [-0.7e7, 0.4e7, 1.5e7]])
    W = np.zeros((9, nx, nz, ny))
    for c, (a, b) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]):
        W[3 + c] = sig[a, b]
    o.put("W", W)
    Ri = np.linalg.inv(R)  # rows: grad xi, grad eta, grad zeta (metric in grid-spacing units)
    J = np.linalg.det(R)
    t2 = J * sig @ Ri[1]
    t3 = J * sig @ Ri[2]
    FW = np.zeros((9, 2, nz, ny))
    for r in range(3):
        FW[3 + r] = t2[r]
        FW[6 + r] = t3[r]
    o.put("FW", FW)
    M = o.get("M").reshape(13, nx, nz, ny)
    np.testing.assert_allclose(M[9][3:-3, 3:-3, 3:-3], J, rtol=1e-10)
    o.smooth()
    FWs = o.get("FWs").reshape(9, 2, nz, ny)[:, :, 3:-3, 3:-3]
    for r in range(3):
        np.testing.assert_allclose(FWs[3 + r], t2[r], rtol=1e-10, atol=1e-6 * abs(t2).max())
        np.testing.assert_allclose(FWs[6 + r], t3[r], rtol=1e-10, atol=1e-6 * abs(t3).max())
    FW2 = o.get("FW").reshape(9, 2, nz, ny)[:, :, 3:-3, 3:-3]
    np.testing.assert_allclose(FW2[3:], FW[3:, :, 3:-3, 3:-3], rtol=1e-10, atol=1e-6 * abs(t2).max())


def test_smooth_sides_read_their_own_planes():
    """The minus-side average reads planes i0-1..-3 only and the plus side
    i0+1..+3 only (smooth.cu:198, 268): changing the stress on one side moves
    only that side's smoothed tractions."""
    o = _smooth_oracle(np.eye(3))
    d = o.dims()
    nx, ny, nz, i0 = d["nx"], d["ny"], d["nz"], d["i0"]
    rng = np.random.default_rng(3)
    W = rng.standard_normal((9, nx, nz, ny)) * 1e6
    o.put("W", W)
    o.put("FW", rng.standard_normal((9, 2, nz, ny)) * 1e5)
    fw0 = o.get("FW")
    o.smooth()
    base = o.get("FWs").reshape(9, 2, nz, ny).copy()
    W2 = W.copy()
    W2[3:, i0 + 1:] += 5e6  # perturb the plus side only
    o.put("W", W2)
    o.put("FW", fw0)
    o.smooth()
    new = o.get("FWs").reshape(9, 2, nz, ny)
    np.testing.assert_array_equal(new[:, 0], base[:, 0])
    assert np.abs(new[3:, 1, 3:-3, 3:-3] - base[3:, 1, 3:-3, 3:-3]).max() > 1.0


# ---------------------------------------------------------------------------
# Thermal pressurization (thermpress.cu:22-233): temperature T and pore
# pressure P on a 1-D grid x = i dh normal to the fault, mirror symmetric at
# both ends, T' = a_th T'' + omega(x), P' = a_hy P'' + Lambda T', with the
# frictional heating tau V spread as a Gaussian of half-width h.
TP_LAMBDA, TP_RHOC, TP_H, TP_ATH = 0.1e6, 2.7e6, 20e-3, 1.0e-6


def _tp_oracle(TP_n, hy, DT=0.008):
    case = cases.rs_case(NX=12, NY=64, NZ=36, Friction_type=3, TP_n=TP_n, DT=DT)
    case["stress"]["TP_hy"] = np.full((36, 64), hy)
    o = Oracle(case["params"])
    cases.apply_to_oracle(case, o)
    o.setup()
    n2 = 64 * 36
    united = o.get("united").astype(bool) if o.size("united") else np.zeros(n2, bool)
    return o, n2, united


def _tp_drive(o, n2, V=1e-3, f=0.6, tn=-100e6):
    o.put("Vs1", np.full(n2, V))
    o.put("Vs2", np.zeros(n2))
    o.put("friction", np.full(n2, f))
    o.put("Tn", np.full(n2, tn))


def test_thermpress_undrained_limit_p_equals_lambda_t():
    """Without hydraulic diffusion (a_hy = 0) P' = Lambda T', so from zero
    P = Lambda T at every level, whatever the heating and the thermal
    diffusion (the coupling takes all of T', diffusive part included)."""
    o, n2, united = _tp_oracle(21, 0.0)
    _tp_drive(o, n2)
    for _ in range(10):
        o.thermpress()
    T = o.get("TP_T").reshape(21, n2)[:, ~united]
    P = o.get("TP_P").reshape(21, n2)[:, ~united]
    assert T[0].min() > 0
    np.testing.assert_allclose(P, TP_LAMBDA * T, rtol=1e-12, atol=1e-9 * np.abs(P).max())


def test_thermpress_mirror_boundaries_conserve_heat():
    """The mirror-symmetric discrete Laplacian conserves the trapezoid
    integral of T over [0, w]: after each call it has grown by exactly the
    integral of the heating over the substeps (constant here: P stays small
    against the normal stress only approximately, so tau is recomputed from
    the oracle's own P(0) per substep)."""
    TP_n, hy, DT = 31, 4e-4, 0.008
    o, n2, united = _tp_oracle(TP_n, hy, DT)
    V, f, tn = 1e-2, 0.6, -100e6
    _tp_drive(o, n2, V, f, tn)
    w = math.sqrt(4.0 * max(TP_ATH, hy) * 25.0)
    dh = w / (TP_n - 1)
    nsub = int(DT / (0.1 * dh * dh / max(TP_ATH, hy))) + 1
    dt = DT / nsub
    x = np.arange(TP_n) * dh
    g = np.exp(-0.5 * x * x / TP_H ** 2) / (TP_RHOC * TP_H * math.sqrt(2 * math.pi))
    wts = np.full(TP_n, dh)
    wts[0] = wts[-1] = dh / 2
    q = int(np.flatnonzero(~united)[0])
    heat = 0.0
    for _ in range(3):
        T0 = o.get("TP_T").reshape(TP_n, n2)[:, q].copy()
        P0 = o.get("TP_P").reshape(TP_n, n2)[:, q].copy()
        # replay the substeps' heating (tau from P(0) of each substep) with
        # an independent implementation of the same explicit update
        T, P = T0.copy(), P0.copy()
        for _s in range(nsub):
            tauV = -f * (tn + P[0]) * V
            om = g * tauV
            lap = lambda u: (np.r_[u[1:], u[-2]] - 2 * u + np.r_[u[1], u[:-1]]) / dh ** 2
            dT = TP_ATH * lap(T) + om
            dP = hy * lap(P) + TP_LAMBDA * dT
            heat += dt * float(wts @ om)
            T, P = T + dt * dT, P + dt * dP
        o.thermpress()
        T1 = o.get("TP_T").reshape(TP_n, n2)[:, q]
        assert abs(float(wts @ T1) - heat) <= 1e-11 * heat
        np.testing.assert_allclose(T1, T, rtol=1e-12)


def test_thermpress_matches_continuous_solution():
    """Heating at a rate that stays constant (P << |tau_n|), on a domain
    wide against the diffusion length (a_hy = 1e-2 m^2/s: w = 1 m), with
    T nearly adiabatic (a_th = 1e-6): the fault pressure follows the
    continuous solution of P' = a_hy P'' + Lambda omega(x) for a Gaussian
    source of half-width h,
        P(0, t) = Lambda (tau V / rho c) (sqrt(h^2 + 2 a_hy t) - h) / (a_hy sqrt(2 pi)),
    and T(0, t) the adiabatic (tau V / rho c) t / (h sqrt(2 pi)).  TP_n = 201
    (dh = 5 mm = h / 4); bound 1.5% (grid, explicit Euler, the small
    a_th)."""
    hy, V, f, tn = 1e-2, 1e-3, 0.6, -100e6
    o, n2, united = _tp_oracle(201, hy, DT=0.01)
    _tp_drive(o, n2, V, f, tn)
    steps = 50
    for _ in range(steps):
        o.thermpress()
    t = steps * 0.01
    q = int(np.flatnonzero(~united)[0])
    P0 = o.get("TP_P").reshape(201, n2)[0, q]
    T0 = o.get("TP_T").reshape(201, n2)[0, q]
    tauV = -f * tn * V
    P_ref = TP_LAMBDA * tauV / TP_RHOC * (math.sqrt(TP_H ** 2 + 2 * hy * t) - TP_H) / (hy * math.sqrt(2 * math.pi))
    T_ref = tauV / TP_RHOC * t / (TP_H * math.sqrt(2 * math.pi))
    assert P0 < 1e-3 * abs(tn)
    assert abs(P0 / P_ref - 1) < 0.015, (P0, P_ref)
    assert abs(T0 / T_ref - 1) < 0.015, (T0, T_ref)
