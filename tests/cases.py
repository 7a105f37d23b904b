"""Small seeded problem definitions shared by the oracle and HIP tests.

Each case is a dict with 'params' (params.json keys) and optional input
arrays ('geom': fault surface x,y,z [NZ][NY]; 'media1d'; 'stress': init
stress variables [NZ][NY]).  apply_to_solver / apply_to_oracle feed the same
inputs to both sides.
"""
import numpy as np


def base_params(NX=24, NY=40, NZ=32, DH=100.0, DT=0.008, **kw):
    p = {
        "TMAX": 1.0, "DT": DT, "DH": DH, "NX": NX, "NY": NY, "NZ": NZ,
        "PX": 1, "PY": 1, "PZ": 1, "INPORT_GRID_TYPE": 0, "INPORT_STRESS_TYPE": 0,
        "Friction_type": 0, "EXPORT_TIME_SKIP": 1, "PML_N": 8, "PML_velocity": 6000.0,
        "PML_bmax": 3.0, "PML_fc": 1.0, "DAMP_N": 8,
        "bi_vp1": 6000.0, "bi_vs1": 3464.0, "bi_rho1": 2670.0,
        "bi_vp2": 6000.0, "bi_vs2": 3464.0, "bi_rho2": 2670.0,
        "usePML": 1, "useCerjan": 0, "FreeSurface": 1, "NormalRefine": 1, "Rupture": 0,
    }
    p.update(kw)
    return p


def bent_geometry(NY, NZ, DH, amp=0.3, seed=0):
    """Fault surface x(j,k), y, z [NZ][NY]: a smooth seeded bend along strike
    and dip; z is 0 at the top row (k = NZ-1) as the reference's grid."""
    rng = np.random.default_rng(seed)
    j = np.arange(NY)[None, :]
    k = np.arange(NZ)[:, None]
    ph = rng.uniform(0, 2 * np.pi, 2)
    x = amp * DH * (np.sin(2 * np.pi * j / max(NY / 2, 1) + ph[0]) * np.cos(np.pi * k / NZ + ph[1]))
    y = (j - NY // 2) * DH + 0.0 * k
    z = (k - NZ + 1) * DH + 0.0 * j
    return x.astype(np.float64), np.broadcast_to(y, (NZ, NY)).astype(np.float64), np.broadcast_to(z, (NZ, NY)).astype(np.float64)


def volume_case(curved=True, **kw):
    p = base_params(**kw)
    case = {"params": p}
    if curved:
        p["INPORT_GRID_TYPE"] = 1
        case["geom"] = bent_geometry(p["NY"], p["NZ"], p["DH"])
    return case


def sw_case(NX=30, NY=80, NZ=50, curved=False, seed=1, **kw):
    """Slip-weakening rupture on the x-mid plane with a nucleation patch."""
    p = base_params(NX=NX, NY=NY, NZ=NZ, Rupture=1, Friction_type=0, INPORT_STRESS_TYPE=1,
                    Fault_grid=[31, NY - 30, 31, NZ], Asp_grid=[0, 0, 0, 0],
                    mu_s=0.677, mu_d=0.525, Dc=0.4, C0=0.0, smooth_load_T=0.0, **kw)
    case = {"params": p}
    if curved:
        p["INPORT_GRID_TYPE"] = 1
        case["geom"] = bent_geometry(NY, NZ, p["DH"], amp=0.2, seed=seed)
    rng = np.random.default_rng(seed)
    jj, kk = np.meshgrid(np.arange(NY), np.arange(NZ))
    tn = -120e6 * np.ones((NZ, NY))
    ts = -70e6 + 1e6 * rng.standard_normal((NZ, NY))
    jc, kc = (31 + NY - 30) // 2, NZ - 10
    patch = (np.abs(jj - jc) <= 3) & (np.abs(kk - kc) <= 3)
    ts[patch] = -0.677 * 120e6 * 1.01
    mu_s = np.where((jj >= 30) & (jj < NY - 30) & (kk >= 30), 0.677, 10000.0)
    case["stress"] = {
        "Tx": tn, "Ty": ts, "Tz": np.zeros((NZ, NY)),
        "dTx": np.zeros((NZ, NY)), "dTy": np.zeros((NZ, NY)), "dTz": np.zeros((NZ, NY)),
        "mu_s": mu_s, "mu_d": 0.525 * np.ones((NZ, NY)), "Dc": 0.4 * np.ones((NZ, NY)), "C0": np.zeros((NZ, NY)),
    }
    return case


def rs_case(NX=30, NY=80, NZ=50, curved=False, seed=2, **kw):
    """Rate-and-state aging law with a Gaussian-in-time overstress patch."""
    p = base_params(NX=NX, NY=NY, NZ=NZ, Rupture=1, INPORT_STRESS_TYPE=1,
                    Fault_grid=[31, NY - 30, 31, NZ], Asp_grid=[0, 0, 0, 0],
                    RS_V0=1e-6, RS_f0=0.6, RS_Vini=1e-12, RS_fw=0.2, smooth_load_T=0.1,
                    **dict({"RS_L": 0.02, "Friction_type": 1}, **kw))
    case = {"params": p}
    if curved:
        p["INPORT_GRID_TYPE"] = 1
        case["geom"] = bent_geometry(NY, NZ, p["DH"], amp=0.2, seed=seed)
    rng = np.random.default_rng(seed)
    jj, kk = np.meshgrid(np.arange(NY), np.arange(NZ))
    a = 0.008 + 0.008 * rng.uniform(0, 1, (NZ, NY))
    b = 0.012 * np.ones((NZ, NY))
    L = 0.02 * np.ones((NZ, NY))
    V0, f0, Vini = 1e-6, 0.6, 1e-12
    sn, ts0 = 120e6, 75e6
    # steady state for the initial slip rate: psi = a ln(2 V0/V sinh(tau/(a sn)))
    psi = a * np.log(2 * V0 / Vini * np.sinh(ts0 / (a * sn)))
    jc, kc = (31 + NY - 30) // 2, NZ - 10
    r = np.hypot(jj - jc, kk - kc) * p["DH"]
    R = 800.0
    F = np.where(r < R, np.exp(np.minimum(r * r / (r * r - R * R + 1e-9), 0)), 0.0)
    case["stress"] = {
        "Tx": -sn * np.ones((NZ, NY)), "Ty": -ts0 * np.ones((NZ, NY)), "Tz": np.zeros((NZ, NY)),
        "dTx": np.zeros((NZ, NY)), "dTy": -25e6 * F, "dTz": np.zeros((NZ, NY)),
        "a": a, "b": b, "L": L, "Vw": 0.1 * np.ones((NZ, NY)), "State": psi,
    }
    return case


def apply_to_solver(case, solver):
    if "geom" in case:
        solver.set_fault_geometry(*case["geom"])
    if "media1d" in case:
        solver.set_media1d(*case["media1d"])
    if case.get("topo") is not None:
        solver.set_topography(case["topo"])
    for k, v in case.get("stress", {}).items():
        solver.set_init_stress(k, v)


def apply_to_oracle(case, orc):
    if "geom" in case:
        orc.set_geometry(*case["geom"])
    if "media1d" in case:
        orc.set_media1d(*case["media1d"])
    if case.get("topo") is not None:
        orc.set_topography(case["topo"])
    for k, v in case.get("stress", {}).items():
        orc.set_stress(k, v)


def random_wavefield(nx, ny, nz, seed=0, scale=(1.0, 1e6)):
    """Smooth-ish random interior wavefield (velocities ~1, stresses ~1e6)
    with zero ghost layers, layout [9][nx][nz][ny]."""
    rng = np.random.default_rng(seed)
    W = np.zeros((9, nx, nz, ny))
    core = rng.standard_normal((9, nx - 6, nz - 6, ny - 6))
    # mild smoothing so the stencil sees structured data
    for ax in (1, 2, 3):
        core = 0.5 * core + 0.25 * (np.roll(core, 1, ax) + np.roll(core, -1, ax))
    W[:3, 3:-3, 3:-3, 3:-3] = core[:3] * scale[0]
    W[3:, 3:-3, 3:-3, 3:-3] = core[3:] * scale[1]
    return W
