"""FP64 restatement of the learned Newton-guess network (4-64-64-1, ReLU),
its loss gradient, and the exact rate-and-state solve it is trained to
predict.  Test infrastructure only: the tests use it as the checker for the
GPU forward, gradient and training kernels; the product path never does.

The rate-state trial (trial_ratestate.cu:686-702, NRsearch 20-57) solves
    2 V0 e^(-psi/a) sinh x + coef x - Vt = 0,   V = 2 V0 e^(-psi/a) sinh x.
With e = ln(2 V0/coef) - psi/a this is e^e sinh x + x = Vt/coef, and with
the features f0 = T(e + Vt/coef) (T(s) = s for s <= 0, ln(1 + s) above),
f1 = Vt/coef, f2 = log10 coef (f3 = log10 of the previous slip rate does not
enter) log10 V = f2 + (e + ln sinh x)/ln 10.

Parameter blob (float32, 4555 values):
  xmean[4] xscale[4] ymean yscale W1[64][4] b1[64] W2[64][64] b2[64] w3[64] b3
Forward: x = (f - xmean) * xscale; h1 = relu(W1 x + b1); h2 = relu(W2 h1 + b2);
o = w3 . h2 + b3; output o * yscale + ymean.  With bf16=True the roundings of
the GPU kernels are applied (W2 and h1 rounded to bfloat16, round to nearest
even); the gradient treats those roundings as identity, as the GPU does.
Loss: mean_s (o_s - (t_s - ymean)/yscale)^2.

Parity pins (tests/test_mlp_oracle.py): the bf16 rounding against torch's
float32 -> bfloat16 conversion; the forward pass against a network built to
compute a known function (asymmetric in W1 / W2 orientation); the gradient
against central finite differences; the sample solve against scipy's brentq.
"""
import numpy as np

N_IN, N_HID, N_PARAMS = 4, 64, 4555
OFF = {"xmean": 0, "xscale": 4, "ymean": 8, "yscale": 9, "W1": 10, "b1": 266, "W2": 330, "b2": 4426,
       "w3": 4490, "b3": 4554}


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, np.float64)
    return {
        "xmean": P[0:4], "xscale": P[4:8], "ymean": P[8], "yscale": P[9],
        "W1": P[OFF["W1"]:OFF["b1"]].reshape(N_HID, N_IN), "b1": P[OFF["b1"]:OFF["W2"]],
        "W2": P[OFF["W2"]:OFF["b2"]].reshape(N_HID, N_HID), "b2": P[OFF["b2"]:OFF["w3"]],
        "w3": P[OFF["w3"]:OFF["b3"]], "b3": P[OFF["b3"]],
    }


def pack(d):
    P = np.zeros(N_PARAMS, np.float64)
    P[0:4], P[4:8], P[8], P[9] = d["xmean"], d["xscale"], d["ymean"], d["yscale"]
    P[OFF["W1"]:OFF["b1"]] = np.ravel(d["W1"])
    P[OFF["b1"]:OFF["W2"]] = d["b1"]
    P[OFF["W2"]:OFF["b2"]] = np.ravel(d["W2"])
    P[OFF["b2"]:OFF["w3"]] = d["b2"]
    P[OFF["w3"]:OFF["b3"]] = d["w3"]
    P[OFF["b3"]] = d["b3"]
    return P


def _layers(P, F, bf16):
    d = unpack(P)
    x = (np.asarray(F, np.float64).reshape(-1, N_IN) - d["xmean"]) * d["xscale"]
    if bf16:  # the GPU standardises in float32
        x = ((np.asarray(F, np.float32).reshape(-1, N_IN) - d["xmean"].astype(np.float32))
             * d["xscale"].astype(np.float32)).astype(np.float64)
    z1 = x @ d["W1"].T + d["b1"]
    h1 = np.maximum(z1, 0.0)
    W2 = d["W2"]
    if bf16:
        h1 = bf16_round(h1)
        W2 = bf16_round(W2)
    z2 = h1 @ W2.T + d["b2"]
    h2 = np.maximum(z2, 0.0)
    o = h2 @ d["w3"] + d["b3"]
    return d, x, z1, h1, W2, z2, h2, o


def forward(P, F, bf16=True):
    """Network output (log10 slip rate) for feature rows F [n][4]."""
    d, *_, o = _layers(P, F, bf16)
    return o * d["yscale"] + d["ymean"]


def loss_grad(P, F, T, bf16=True):
    """(loss, gradient [4555], zero for the standardisation entries)."""
    d, x, z1, h1, W2, z2, h2, o = _layers(P, F, bf16)
    n = len(o)
    tn = (np.asarray(T, np.float64).ravel() - d["ymean"]) / d["yscale"]
    e = o - tn
    loss = float(np.mean(e * e))
    do = 2.0 * e / n
    g = {k: np.zeros_like(np.asarray(v, np.float64)) for k, v in d.items()}
    g["b3"] = do.sum()
    g["w3"] = h2.T @ do
    dz2 = np.outer(do, d["w3"]) * (z2 > 0)
    g["b2"] = dz2.sum(0)
    g["W2"] = dz2.T @ h1
    dz1 = (dz2 @ W2) * (z1 > 0)
    g["b1"] = dz1.sum(0)
    g["W1"] = dz1.T @ x
    for k in ("xmean", "xscale", "ymean", "yscale"):
        g[k] = np.zeros_like(np.asarray(d[k], np.float64))
    return loss, pack(g)


def solve_x(e, r):
    """x* of e^e sinh x + x = r (r > 0, e <= 0) in FP64.  The left side is
    convex and increasing on x > 0, so Newton from a point where it exceeds
    r decreases monotonically to x*.  The start min(r, ln(2 r + 1) - e) is
    such a point (x* <= r, and x* <= asinh(r e^-e) <= ln(2 r + 1) - e)."""
    e = np.asarray(e, np.float64)
    r = np.asarray(r, np.float64)
    x = np.minimum(r, np.log(2.0 * r + 1.0) - e)
    for _ in range(500):
        ep, em = np.exp(e + x), np.exp(e - x)
        g = 0.5 * (ep - em) + x - r
        dg = 0.5 * (ep + em) + 1.0
        xn = x - g / dg
        done = np.abs(xn - x) <= 4e-16 * np.abs(x)
        x = xn
        if np.all(done):
            break
    return x


def log_sinh(x):
    x = np.asarray(x, np.float64)
    return np.where(x > 20, x - np.log(2.0) + np.log1p(-np.exp(-2.0 * np.minimum(x, 700))),
                    np.log(np.sinh(np.maximum(x, 1e-300))))


def feature_s(f0):
    """s = e + Vt/coef from the feature f0 = T(s)."""
    f0 = np.asarray(f0, np.float64)
    return np.where(f0 <= 0.0, f0, np.expm1(np.maximum(f0, 0.0)))


def target_log10V(F):
    """Exact log10 slip rate of feature rows F [n][4] (f3 does not enter)."""
    F = np.asarray(F, np.float64).reshape(-1, N_IN)
    e = feature_s(F[:, 0]) - F[:, 1]
    x = solve_x(e, F[:, 1])
    return F[:, 2] + (e + log_sinh(x)) / np.log(10.0)


def synthetic_samples(n, seed=0):
    """Feature rows spanning the trial states of the rate-state cases, from
    locked (V ~ 1e-30 m/s) to sliding (V ~ 10 m/s) nodes, and their exact
    targets.  Drawn as x* = friction/a in [20, 100] and u = x* + e in
    [-62, 8], which fixes Vt/coef = x* + e^e sinh x*."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(20.0, 100.0, n)
    u = rng.uniform(-62.0, 8.0, n)
    e = u - x
    f1 = x + 0.5 * (np.exp(u) - np.exp(e - x))
    f2 = rng.uniform(-5.0, -1.0, n)           # log10 coef
    s = e + f1
    F = np.stack([np.where(s <= 0.0, s, np.log1p(np.maximum(s, 0.0))), f1, f2, np.zeros(n)], 1)
    F = F.astype(np.float32).astype(np.float64)  # the targets of the stored features
    t = target_log10V(F)
    F[:, 3] = t + rng.normal(0.0, 0.5, n)     # previous slip rate near the new one
    return F.astype(np.float32), t.astype(np.float32)
