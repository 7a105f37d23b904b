"""Seeded rupture cases for the accuracy comparison against the FP64 oracle
(final slip, rupture time, peak slip rate, surface seismograms and on-fault
station time series)."""
import numpy as np

from dynfd3dcuda_amd import configs


def sw_curved():
    """Slip weakening on a bent fault under topography with a 3-layer medium
    and ADE-PML: the ingredients of the curved-grid benchmark, on an 80 x 160
    x 80 grid at 100 m, 400 steps of 0.006 s (the rupture crosses most of the
    fault; the stability limit of this grid is 0.0074 s)."""
    rng = np.random.default_rng(11)
    NX, NY, NZ, DH, DT, nt = 80, 160, 80, 100.0, 0.006, 400
    p = configs._params(NX, NY, NZ, DH, DT, nt, PML_N=12)
    geom = configs.bent_fault_geometry(NY, NZ, DH, 20.0, rng)
    return {"params": p, "nt": nt, "geom": geom,
            "media1d": configs.three_layer_media(),
            "topo": configs.sum_of_gaussians_topography(NX, NY, DH, rng, hmax=500.0),
            "stress": configs._sw_stress(NY, NZ, DH, rng, 70e6, 3e6, 1e3, 3.0,
                                         frame=configs.fault_frame(geom, DH))}


def rs_curved():
    """Rate-and-state aging law with a Gaussian-in-time overstress on the same
    kind of grid, 400 steps.  L = 0.1 m keeps the explicit state update
    stable (DT V / L < 2.8) up to 46 m/s slip rate."""
    c = sw_curved()
    rng = np.random.default_rng(12)
    p = c["params"]
    p.update({"Friction_type": 1, "RS_V0": 1e-6, "RS_f0": 0.6, "RS_Vini": 1e-12, "RS_L": 0.02,
              "RS_fw": 0.2, "smooth_load_T": 0.5})
    c["stress"] = configs._rs_stress(p["NY"], p["NZ"], p["DH"], rng, frame=configs.fault_frame(c["geom"], p["DH"]), L=0.1)
    return c


def tp_curved():
    """The rate-and-state case with thermal pressurization (Friction_type 3:
    slip law with flash weakening, pore pressure from the 1-D diffusion
    normal to the fault, TP_n = 21, hydraulic diffusivity 2^-11 m^2/s, exact
    in FP32), 400 steps.  L = 0.4 m (as in the reference's slip-law set-up,
    rk4.cu:883-886): with L = 0.1 m the slip-law state update of this
    strongly weakening rupture goes unstable near step 200 in the oracle
    itself.  The pore pressure approaches the normal stress, and slip rates
    reach about 90 m/s."""
    c = rs_curved()
    p = c["params"]
    p.update({"Friction_type": 3, "TP_n": 21, "RS_L": 0.4})
    c["stress"]["L"] = np.full_like(c["stress"]["L"], 0.4)
    c["stress"]["TP_hy"] = np.full_like(c["stress"]["L"], 2.0 ** -11)
    return c


CASES = {"sw_curved": sw_curved, "rs_curved": rs_curved, "tp_curved": tp_curved}


def apply_to_oracle(case, o):
    o.set_geometry(*case["geom"])
    o.set_media1d(*case["media1d"])
    o.set_topography(case["topo"])
    for k, v in case["stress"].items():
        o.set_stress(k, v)


def apply_to_solver(case, s):
    configs.apply(case, s)


def receivers(d):
    """Surface points (i, j, k=nk-1) on both sides of the fault."""
    i0 = d["nx"] // 2 - 3
    nj, nk = d["nj"], d["nk"]
    return [(i0 - 5, nj // 2, nk - 1), (i0 + 5, nj // 2, nk - 1), (i0 - 10, nj // 3, nk - 1),
            (i0 + 10, 2 * nj // 3, nk - 1), (i0 - 3, nj // 2 + 10, nk - 1)]


def sample_receivers(W, d, rec):
    W = W.reshape(9, d["nx"], d["nz"], d["ny"])
    return np.array([[W[c, i + 3, k + 3, j + 3] for c in range(3)] for (i, j, k) in rec])


FAULT_QUANTITIES = ("Vs1", "Vs2", "slip", "Ts1", "Ts2", "Tn")


def fault_stations(d):
    """Fault nodes (j, k) of the on-fault time series: the hypocentre of the
    nucleation patch (configs._sw_stress: j = nj/2, k = nk - 1 - (nk - 30)/2),
    four nodes along strike at 25 and 45 cells from it, one up dip near the
    free surface and one down dip."""
    nj, nk = d["nj"], d["nk"]
    jc, kc = nj // 2, nk - 1 - int(round(0.5 * (nk - 30)))
    return [(jc, kc), (jc - 25, kc), (jc + 25, kc), (jc - 45, kc), (jc + 45, kc), (jc, nk - 6), (jc, kc - 18)]


def sample_fault(get, d, st):
    """The FAULT_QUANTITIES at the stations: get(name) returns the [nk][nj]
    fault array of a quantity, flattened."""
    nj = d["nj"]
    vals = [np.asarray(get(q)).reshape(-1) for q in FAULT_QUANTITIES]
    return np.array([[v[k * nj + j] for v in vals] for (j, k) in st])
