"""Seeded synthetic inputs for the benchmark configurations.

Each builder returns a dict with
  params   : params.json keys (the reference's key set plus the runtime flags)
  geom     : fault surface x, y, z [NZ][NY] (Fault_geometry.nc variables)
  media1d  : (depth, rho, vp, vs) layer table (Media1D file contents)
  stress   : Fault_init_stress.nc variables [NZ][NY]
  topo     : surface height h [NX][NY] added to z with a linear taper to the
             bottom (extension: the reference's grid cannot vary z with x)
Grids keep the reference's axes: x is fault-normal (fault on the x-mid
plane), y along strike, z along dip with the free surface on top.
"""
import math

import numpy as np

FAULT_BORDER = 30  # 'united' border of the reference (init.cu:806-816)


def gaussian_random_field(shape, corr_cells, rng):
    """Zero-mean, unit-sd field with a Gaussian correlation of corr_cells."""
    w = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for n in shape]
    kk = np.meshgrid(*k, indexing="ij")
    k2 = sum(x * x for x in kk)
    filt = np.exp(-0.5 * (2 * np.pi) ** 2 * k2 * corr_cells ** 2)
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * filt))
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def smooth_curve(n, corr_cells, rng):
    return gaussian_random_field((n,), corr_cells, rng)


def flat_geometry(NY, NZ, DH):
    j = np.arange(NY)[None, :]
    k = np.arange(NZ)[:, None]
    x = np.zeros((NZ, NY))
    y = np.broadcast_to((j - NY // 2) * DH, (NZ, NY)).astype(np.float64)
    z = np.broadcast_to((k - NZ + 1) * DH, (NZ, NY)).astype(np.float64)
    return x, y.copy(), z.copy()


def bent_fault_geometry(NY, NZ, DH, max_deg, rng):
    """Vertical fault whose strike bends by 0..max_deg along a seeded curve."""
    t = smooth_curve(NY, corr_cells=NY / 6, rng=rng)
    t = (t - t.min()) / max(t.max() - t.min(), 1e-12)
    theta = np.deg2rad(max_deg) * t
    xs = np.concatenate([[0.0], np.cumsum(np.tan(theta[:-1]) * DH)])
    xs -= xs[NY // 2]
    x, y, z = flat_geometry(NY, NZ, DH)
    x[:] = xs[None, :]
    return x, y, z


def fault_frame(geom, DH):
    """Unit normal n and along-strike vector s1 of a vertical fault surface
    x(j) (geometry x, y, z [NZ][NY]); returns n, s1 as [3][NZ][NY]."""
    x = geom[0]
    dxdy = np.gradient(x, axis=1) / DH
    nrm = np.sqrt(1.0 + dxdy * dxdy)
    n = np.stack([1.0 / nrm, -dxdy / nrm, np.zeros_like(x)])
    s1 = np.stack([dxdy / nrm, 1.0 / nrm, np.zeros_like(x)])
    return n, s1


def local_traction(sn, tau, frame):
    """Global traction vector (Tx, Ty, Tz), the form Fault_init_stress takes
    (init.cu:904-1022), of normal stress sn and along-strike shear tau (the
    sign of the reference's flat-fault convention Ty = tau) on the fault
    frame (n, s1).  On a flat fault this is (sn, tau, 0)."""
    n, s1 = frame
    return [sn * n[c] + tau * s1[c] for c in range(3)]


def sum_of_gaussians_topography(NX, NY, DH, rng, n_hills=6, hmax=800.0, wmin=2e3, wmax=6e3):
    xg = (np.arange(NX) - NX // 2) * DH
    yg = (np.arange(NY) - NY // 2) * DH
    X, Y = np.meshgrid(xg, yg, indexing="ij")
    h = np.zeros((NX, NY))
    for _ in range(n_hills):
        cx, cy = rng.uniform(xg[0], xg[-1]), rng.uniform(yg[0], yg[-1])
        w = rng.uniform(wmin, wmax)
        amp = rng.uniform(0.2, 1.0)
        h += amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (w * w))
    if h.max() > 0:
        h *= hmax / h.max()
    return h


def three_layer_media():
    vs = np.array([2000.0, 3000.0, 3464.0])
    depth = np.array([0.0, 1000.0, 3000.0])
    vp = vs * math.sqrt(3.0)
    rho = np.array([2400.0, 2600.0, 2670.0])
    return depth, rho, vp, vs


def _sw_stress(NY, NZ, DH, rng, mean, sd, corr_m, nuc_km, frame=None):
    mu_s, sn = 0.677, 120e6
    tau = mean + sd * gaussian_random_field((NZ, NY), corr_m / DH, rng)
    tau = np.minimum(tau, 0.99 * mu_s * sn)  # clipped below static strength
    jj, kk = np.meshgrid(np.arange(NY), np.arange(NZ))
    jc = NY // 2
    kc = NZ - 1 - int(round(0.5 * (NZ - FAULT_BORDER)))
    half = int(round(nuc_km * 1e3 / DH / 2))
    patch = (np.abs(jj - jc) <= half) & (np.abs(kk - kc) <= half)
    tau[patch] = 1.01 * mu_s * sn
    z = np.zeros((NZ, NY))
    if frame is None:
        frame = (np.stack([1 + z, z, z]), np.stack([z, 1 + z, z]))
    Tx, Ty, Tz = local_traction(-sn * np.ones((NZ, NY)), -tau, frame)
    return {
        "Tx": Tx, "Ty": Ty, "Tz": Tz,
        "dTx": z.copy(), "dTy": z.copy(), "dTz": z.copy(),
        "mu_s": mu_s * np.ones((NZ, NY)), "mu_d": 0.525 * np.ones((NZ, NY)),
        "Dc": 0.4 * np.ones((NZ, NY)), "C0": z.copy(),
    }


def _boxcar(x, W, w):
    """Smooth boxcar of the SCEC rate-state benchmarks: 1 for |x| <= W,
    0.5 (1 + tanh(w/(|x|-W-w) + w/(|x|-W))) for W < |x| < W + w, 0 beyond."""
    x = np.abs(x)
    out = np.zeros_like(x, dtype=np.float64)
    out[x <= W] = 1.0
    m = (x > W) & (x < W + w)
    xm = x[m]
    out[m] = 0.5 * (1.0 + np.tanh(w / (xm - W - w) + w / (xm - W)))
    return out


def _rs_stress(NY, NZ, DH, rng, nuc_km=3.0, frame=None, L=0.02):
    """Aging-law setup in the manner of the SCEC TPV101 benchmark: velocity
    weakening a = 0.008 (b = 0.012) in the middle of the frictional area,
    rising to velocity strengthening a = 0.016 over a band (3 km, at most a
    fifth of the extent) at its lateral, bottom and top edges, with a seeded
    +-3% field on a; sigma_n = 120 MPa, tau = 75 MPa + 2 MPa seeded field,
    steady state at V = 1e-12 m/s, and a 25 MPa overstress patch of radius
    nuc_km ramped in time over smooth_load_T (via dT0).  The explicit RK4
    update of the aging-law state is stable while DT V / L stays below about
    2.8, which bounds the slip rate a run with characteristic slip L can
    carry."""
    sn, tau_mean, tau_sd = 120e6, 75e6, 2e6
    V0, Vini = 1e-6, 1e-12
    jj, kk = np.meshgrid(np.arange(NY), np.arange(NZ))
    ly = (NY - 2 * FAULT_BORDER) * DH           # frictional extent along strike
    lz = (NZ - FAULT_BORDER) * DH               # and down dip from the surface
    wy, wz = min(3e3, 0.2 * ly), min(3e3, 0.2 * lz)
    y = (jj - (NY - 1) / 2.0) * DH
    depth = (NZ - 1 - kk) * DH
    B = _boxcar(y, ly / 2 - wy, wy) * _boxcar(depth - lz / 2, lz / 2 - wz, wz)
    a = 0.008 + 0.008 * (1.0 - B)
    a *= 1.0 + 0.03 * np.tanh(gaussian_random_field((NZ, NY), 2e3 / DH, rng))
    b = 0.012 * np.ones((NZ, NY))
    tau = tau_mean + tau_sd * gaussian_random_field((NZ, NY), 1e3 / DH, rng)
    psi = a * np.log(2 * V0 / Vini * np.sinh(tau / (a * sn)))  # steady state at Vini
    jc = NY // 2
    kc = NZ - 1 - int(round(lz / 2 / DH))
    r = np.hypot(jj - jc, kk - kc) * DH
    R = nuc_km * 1e3
    F = np.where(r < R, np.exp(np.minimum(r * r / np.minimum(r * r - R * R, -1e-9), 0.0)), 0.0)
    z = np.zeros((NZ, NY))
    if frame is None:
        frame = (np.stack([1 + z, z, z]), np.stack([z, 1 + z, z]))
    Tx, Ty, Tz = local_traction(-sn * np.ones((NZ, NY)), -tau, frame)
    dTx, dTy, dTz = local_traction(z, -25e6 * F, frame)
    return {
        "Tx": Tx, "Ty": Ty, "Tz": Tz,
        "dTx": dTx, "dTy": dTy, "dTz": dTz,
        "a": a, "b": b, "L": L * np.ones((NZ, NY)), "Vw": 0.1 * np.ones((NZ, NY)), "State": psi,
    }


def _params(NX, NY, NZ, DH, DT, nt, **kw):
    p = {
        "TMAX": nt * DT + DT * 0.5, "DT": DT, "DH": DH, "NX": NX, "NY": NY, "NZ": NZ,
        "PX": 1, "PY": 1, "PZ": 1, "INPORT_GRID_TYPE": 1, "INPORT_STRESS_TYPE": 1,
        "Friction_type": 0, "EXPORT_TIME_SKIP": 10,
        "EXPORT_WAVE_SLICE_X": NX // 2, "EXPORT_WAVE_SLICE_Y": NY // 2, "EXPORT_WAVE_SLICE_Z": NZ - 1,
        "Fault_grid": [FAULT_BORDER + 1, NY - FAULT_BORDER, FAULT_BORDER + 1, NZ],
        "Asp_grid": [0, 0, 0, 0],
        "mu_s": 0.677, "mu_d": 0.525, "Dc": 0.4, "C0": 0.0,
        "PML_N": 20, "PML_velocity": 6000.0, "PML_bmax": 3.0, "PML_fc": 1.0, "DAMP_N": 20,
        "bi_vp1": 6000.0, "bi_vs1": 3464.0, "bi_rho1": 2670.0,
        "bi_vp2": 6000.0, "bi_vs2": 3464.0, "bi_rho2": 2670.0,
        "usePML": 1, "useCerjan": 0, "FreeSurface": 1, "NormalRefine": 1, "Rupture": 1,
        "smooth_load_T": 0.0,
    }
    p.update(kw)
    return p


def flat_sw_100x200x100(seed=0, nt=1500):
    """Plain method, small: NX=100 (normal) x NY=200 (strike) x NZ=100 (dip)
    at 100 m (2.0 M points), flat surface with identity metric, homogeneous
    medium, slip weakening, Cerjan sponge, DT 0.008 s."""
    rng = np.random.default_rng(seed)
    NX, NY, NZ, DH, DT = 100, 200, 100, 100.0, 0.008
    p = _params(NX, NY, NZ, DH, DT, nt, usePML=0, useCerjan=1, NormalRefine=0)
    return {"name": "flat_sw_100x200x100", "params": p, "geom": flat_geometry(NY, NZ, DH),
            "stress": _sw_stress(NY, NZ, DH, rng, 70e6, 3e6, 1e3, 3.0)}


def curved_sw_400x300x200(seed=1, nt=4000, scale=1):
    """Curved grid: 400x300x200 at 50 m, topography, bent fault, 3-layer
    medium, slip weakening, 20-layer ADE-PML, FP32."""
    rng = np.random.default_rng(seed)
    NX, NY, NZ, DH, DT = 400 * scale, 300 * scale, 200 * scale, 50.0 / scale, 0.003 / scale
    p = _params(NX, NY, NZ, DH, DT, nt)
    geom = bent_fault_geometry(NY, NZ, DH, 20.0, rng)
    return {"name": "curved_sw_400x300x200", "params": p, "geom": geom,
            "media1d": three_layer_media(),
            "topo": sum_of_gaussians_topography(NX, NY, DH, rng),
            "stress": _sw_stress(NY, NZ, DH, rng, 70e6, 3e6, 1e3, 3.0, frame=fault_frame(geom, DH))}


def curved_rs_400x300x200(seed=2, nt=4000, scale=1):
    """Rate-and-state aging law with plain Newton on the curved 400x300x200 grid."""
    c = curved_sw_400x300x200(seed=seed, nt=nt, scale=scale)
    rng = np.random.default_rng(seed + 100)
    p = c["params"]
    p.update({"Friction_type": 1, "RS_V0": 1e-6, "RS_f0": 0.6, "RS_Vini": 1e-12, "RS_L": 0.02,
              "RS_fw": 0.2, "smooth_load_T": 1.0})
    c["stress"] = _rs_stress(p["NY"], p["NZ"], p["DH"], rng, frame=fault_frame(c["geom"], p["DH"]))
    c["name"] = "curved_rs_400x300x200"
    return c


def curved_rs_mlp_400x300x200(seed=2, nt=4000, scale=1, mlp_path=None):
    """The curved-grid rate-and-state workload with the learned Newton initial guess."""
    c = curved_rs_400x300x200(seed=seed, nt=nt, scale=scale)
    c["params"].update({"NewtonMode": 2, "NewtonTol": 1e-10, "NewtonFallbackIter": 20})
    if mlp_path:
        c["params"]["NewtonMLP"] = mlp_path
    c["name"] = "curved_rs_mlp_400x300x200"
    return c


def large_rs_mlp_bf16_1600x1000x500(seed=4, nt=8000, sub_nx=None, mlp_path=None):
    """Large, mixed precision: 1600x1000x500 at 25 m, rate-state + learned
    guess, the metric stored as bfloat16 deviations with FP32 media
    (MetricBF16=2).

    sub_nx: keep only the sub_nx fault-normal planes around the fault
    (NX = sub_nx; the same fault, stress, media and the matching rows of the
    same topography): the sub-slab on which the FP64 oracle can run, e.g.
    100x1000x500 (50 M of the 800 M points)."""
    rng = np.random.default_rng(seed)
    NX, NY, NZ, DH, DT = 1600, 1000, 500, 25.0, 0.0015
    geom = bent_fault_geometry(NY, NZ, DH, 20.0, rng)
    topo = sum_of_gaussians_topography(NX, NY, DH, rng)
    stress = _rs_stress(NY, NZ, DH, rng, frame=fault_frame(geom, DH))
    name = "large_rs_mlp_bf16_1600x1000x500"
    if sub_nx is not None:
        if not (0 < sub_nx <= NX and sub_nx % 2 == 0):
            raise ValueError("sub_nx must be an even number of planes in 2..1600")
        lo = NX // 2 - sub_nx // 2
        topo = topo[lo:lo + sub_nx].copy()
        NX = sub_nx
        name = f"{name}_sub{sub_nx}x{NY}x{NZ}"
    p = _params(NX, NY, NZ, DH, DT, nt, Friction_type=1, RS_V0=1e-6, RS_f0=0.6, RS_Vini=1e-12,
                RS_L=0.02, RS_fw=0.2, smooth_load_T=1.0, MetricBF16=2, NewtonMode=2,
                NewtonTol=1e-10, NewtonFallbackIter=20)
    if mlp_path:
        p["NewtonMLP"] = mlp_path
    return {"name": name, "params": p, "geom": geom, "media1d": three_layer_media(), "topo": topo,
            "stress": stress}


WORKLOADS = {f.__name__: f for f in (flat_sw_100x200x100, curved_sw_400x300x200, curved_rs_400x300x200,
                                      curved_rs_mlp_400x300x200, large_rs_mlp_bf16_1600x1000x500)}


def apply(case, solver):
    """Feed a generated case to a Solver (before setup)."""
    if "geom" in case:
        solver.set_fault_geometry(*case["geom"])
    if "media1d" in case:
        solver.set_media1d(*case["media1d"])
    if "topo" in case and case["topo"] is not None:
        solver.set_topography(case["topo"])
    for k, v in case.get("stress", {}).items():
        solver.set_init_stress(k, v)
