"""Writes the FP64-oracle reference outputs of the accuracy cases
(tests/golden/accuracy_<case>.npz) that tests/test_rupture_accuracy.py compares the
HIP solver with: final slip, rupture time (init_t0), peak slip rate,
surface velocity seismograms at a few receivers and time series of slip
rate, slip and tractions at a few fault stations.

    python tests/golden/make_accuracy_fixture.py [case ...]

The cases are defined in tests/accuracy_cases.py (seeded, small enough for the
oracle to finish in a few minutes on 8 cores)."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import accuracy_cases as ac  # noqa: E402
from oracle.binding import Oracle  # noqa: E402


def run(name):
    case = ac.CASES[name]()
    p = case["params"]
    o = Oracle(p)
    ac.apply_to_oracle(case, o)
    o.setup()
    d = o.dims()
    rec = ac.receivers(d)
    st = ac.fault_stations(d)
    nt = case["nt"]
    seis = np.zeros((nt, len(rec), 3))
    fault = np.zeros((nt, len(st), len(ac.FAULT_QUANTITIES)))
    peak = np.zeros(d["nj"] * d["nk"])
    t0 = time.time()
    for it in range(nt):
        o.step(it)
        W = o.get("W")
        seis[it] = ac.sample_receivers(W, d, rec)
        fault[it] = ac.sample_fault(o.get, d, st)
        v = np.hypot(o.get("Vs1"), o.get("Vs2"))
        peak = np.maximum(peak, v)
        if it % 50 == 0:
            print(name, it, f"{time.time() - t0:.1f}s", flush=True)
    out = os.path.join(HERE, f"accuracy_{name}.npz")
    np.savez_compressed(out, slip=o.get("slip"), init_t0=o.get("init_t0"), peak_rate=peak, seis=seis,
                        receivers=np.array(rec), fault_series=fault, fault_stations=np.array(st), nt=nt)
    print("wrote", out, f"{time.time() - t0:.1f}s")


if __name__ == "__main__":
    for n in (sys.argv[1:] or list(ac.CASES)):
        run(n)
