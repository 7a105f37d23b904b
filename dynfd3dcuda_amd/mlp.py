"""Learned initial guess of the rate-and-state Newton solve: parameter
initialisation, GPU training / evaluation through the C ABI, and collection
of training samples from a running solver.

The network (4-64-64-1, ReLU) and its float32 parameter blob are defined in
csrc/mlp.h; the training and evaluation kernels are in csrc/mlp_train.hip.
Samples are what the plain Newton solve produces in each rate-state trial
(trial_ratestate.cu:686-702): the four guess features of every fault node and
the log10 slip rate it solved for (solver field "RS_samples").

    python -m dynfd3dcuda_amd.mlp --workload curved_rs_400x300x200 --steps 400 --out weights.bin
"""
import argparse
import ctypes
import json
import time

import numpy as np

from . import CgfdmError, load_library

N_PARAMS = 4555
N_IN, N_HID = 4, 64
OFF = {"xmean": 0, "xscale": 4, "ymean": 8, "yscale": 9, "W1": 10, "b1": 266, "W2": 330, "b2": 4426,
       "w3": 4490, "b3": 4554}


def _check(lib, rc):
    if rc != 0:
        raise CgfdmError(rc, lib.cgfdm_last_error().decode())


def init_params(feat, target, seed=0):
    """Standardisation from the samples and He-uniform weights (seeded)."""
    feat = np.asarray(feat, np.float64).reshape(-1, N_IN)
    target = np.asarray(target, np.float64).ravel()
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float64)
    mu, sd = feat.mean(0), feat.std(0)
    p[0:4] = mu
    p[4:8] = 1.0 / np.where(sd > 0, sd, 1.0)
    p[8] = target.mean()
    p[9] = max(target.std(), 1e-6)
    lim1, lim2 = np.sqrt(6.0 / N_IN), np.sqrt(6.0 / N_HID)
    p[OFF["W1"]:OFF["b1"]] = rng.uniform(-lim1, lim1, N_HID * N_IN)
    p[OFF["W2"]:OFF["b2"]] = rng.uniform(-lim2, lim2, N_HID * N_HID)
    p[OFF["w3"]:OFF["b3"]] = rng.uniform(-lim2, lim2, N_HID) * 0.5
    return p.astype(np.float32)


def forward(params, feat, mode=0):
    """Network output (log10 slip rate) on the GPU; mode 0 matrix-core, 1 scalar."""
    lib = load_library()
    p = np.ascontiguousarray(params, np.float32)
    f = np.ascontiguousarray(feat, np.float32).reshape(-1, N_IN)
    out = np.empty(len(f), np.float32)
    _check(lib, lib.cgfdm_mlp_forward(p.ctypes.data, f.ctypes.data, len(f), out.ctypes.data, mode))
    return out


def grad(params, feat, target):
    """(loss, gradient) of the mean squared standardised error on the GPU."""
    lib = load_library()
    p = np.ascontiguousarray(params, np.float32)
    f = np.ascontiguousarray(feat, np.float32).reshape(-1, N_IN)
    t = np.ascontiguousarray(target, np.float32).ravel()
    g = np.empty(N_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    _check(lib, lib.cgfdm_mlp_grad(p.ctypes.data, f.ctypes.data, t.ctypes.data, len(f), g.ctypes.data,
                                   loss.ctypes.data))
    return float(loss[0]), g


def train(feat, target, params=None, epochs=30, batch=4096, lr=2e-3, seed=0):
    """Train on the GPU (Adam); returns (params, per-epoch loss)."""
    lib = load_library()
    f = np.ascontiguousarray(feat, np.float32).reshape(-1, N_IN)
    t = np.ascontiguousarray(target, np.float32).ravel()
    p = np.array(init_params(f, t, seed) if params is None else params, np.float32)
    hist = np.zeros(epochs, np.float32)
    _check(lib, lib.cgfdm_mlp_train(p.ctypes.data, f.ctypes.data, t.ctypes.data, len(f), epochs, batch,
                                    ctypes.c_float(lr), seed, hist.ctypes.data))
    return p, hist


def collect(solver, it0, nsteps, every=1, keep=None):
    """Advance the solver nsteps steps and return the rate-state samples of
    the last stage of every `every`-th step (features [n][4], log10 V [n]),
    randomly thinned to at most `keep` samples."""
    d = solver.dims()
    n2 = d["nj"] * d["nk"]
    ok = solver.get("united") == 0
    feats, targs = [], []
    for it in range(it0, it0 + nsteps):
        solver.step(it)
        if (it - it0) % every == 0:
            s = solver.get("RS_samples").reshape(5, n2)
            feats.append(s[:4, ok].T.copy())
            targs.append(s[4, ok].copy())
    f, t = np.concatenate(feats), np.concatenate(targs)
    good = np.all(np.isfinite(f), axis=1) & np.isfinite(t)
    f, t = f[good], t[good]
    if keep is not None and len(t) > keep:
        idx = np.random.default_rng(0).choice(len(t), keep, replace=False)
        f, t = f[idx], t[idx]
    return f, t


def save(params, path):
    np.asarray(params, "<f4").tofile(path)


def load(path):
    p = np.fromfile(path, "<f4")
    if p.size != N_PARAMS:
        raise ValueError(f"{path}: {p.size} values, expected {N_PARAMS}")
    return p


def main(argv=None):
    from . import Solver, configs
    ap = argparse.ArgumentParser(description="collect rate-state samples with plain Newton and train the guess network")
    ap.add_argument("--workload", default="curved_rs_400x300x200",
                    help="rate-state workload of dynfd3dcuda_amd.configs.WORKLOADS")
    ap.add_argument("--seeds", type=int, nargs="+", default=[0])
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--every", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    F, T = [], []
    for seed in a.seeds:
        case = configs.WORKLOADS[a.workload](seed=seed)
        s = Solver(case["params"], NewtonMode=1)
        configs.apply(case, s)
        s.setup()
        t0 = time.time()
        f, t = collect(s, 0, a.steps, a.every)
        print(json.dumps({"seed": seed, "samples": len(t), "seconds": round(time.time() - t0, 1)}), flush=True)
        F.append(f)
        T.append(t)
        s.close()
    f, t = np.concatenate(F), np.concatenate(T)
    t0 = time.time()
    p, hist = train(f, t, epochs=a.epochs)
    print(json.dumps({"samples": len(t), "train_seconds": round(time.time() - t0, 1),
                      "loss_first": float(hist[0]), "loss_last": float(hist[-1])}), flush=True)
    save(p, a.out)


if __name__ == "__main__":
    main()
