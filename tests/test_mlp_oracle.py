"""Pins of the FP64 oracle of the learned Newton guess (oracle/mlp_oracle.py)
against facts fixed independently of it: torch's float32 -> bfloat16
rounding, a network constructed to compute a known function, central finite
differences of the loss, and scipy's bracketing root finder."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 1000), rng.normal(0, 1e4, 1000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 1.0029296875, 3.0e38, 1e-40]]).astype(np.float32)
    # values exactly half-way between two bf16 numbers round to the even one
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    a = np.concatenate([a, ties, -ties])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.bf16_round(a), ref)


def _known_network():
    """Network computing y = 3 x0 - 2 x1 + 0.5 relu(x2 - x3) with x the
    standardised features; W1 and W2 are not symmetric, so a transposed
    operand gives a different function."""
    d = {k: np.zeros_like(v) for k, v in mo.unpack(np.zeros(mo.N_PARAMS)).items()}
    d["xmean"] = np.array([1.0, -2.0, 0.5, 0.25])
    d["xscale"] = np.array([0.5, 2.0, 1.0, 4.0])
    d["ymean"], d["yscale"] = 10.0, 2.0
    W1 = np.zeros((64, 4))
    W1[0, 0], W1[1, 0] = 1.0, -1.0          # relu(x0), relu(-x0)
    W1[2, 1], W1[3, 1] = 1.0, -1.0          # relu(x1), relu(-x1)
    W1[5, 2], W1[5, 3] = 1.0, -1.0          # relu(x2 - x3)
    W2 = np.zeros((64, 64))
    W2[10, 0], W2[11, 1] = 1.0, 1.0         # hidden-2 unit 10 = relu(x0), 11 = relu(-x0)
    W2[20, 2], W2[21, 3] = 1.0, 1.0
    W2[30, 5] = 1.0
    w3 = np.zeros(64)
    w3[10], w3[11], w3[20], w3[21], w3[30] = 3.0, -3.0, -2.0, 2.0, 0.5
    d.update(W1=W1, W2=W2, w3=w3, b1=np.zeros(64), b2=np.zeros(64), b3=0.0)
    return mo.pack(d), d


def test_forward_known_function():
    P, d = _known_network()
    rng = np.random.default_rng(1)
    F = rng.normal(0, 3, (500, 4))
    x = (F - d["xmean"]) * d["xscale"]
    want = (3 * x[:, 0] - 2 * x[:, 1] + 0.5 * np.maximum(x[:, 2] - x[:, 3], 0)) * d["yscale"] + d["ymean"]
    np.testing.assert_allclose(mo.forward(P, F, bf16=False), want, rtol=0, atol=1e-12)
    # transposing W2 in the blob breaks it
    d2 = dict(d)
    d2["W2"] = d["W2"].T.copy()
    assert np.max(np.abs(mo.forward(mo.pack(d2), F, bf16=False) - want)) > 1.0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    F, T = mo.synthetic_samples(300, seed=3)
    F, T = F.astype(np.float64), T.astype(np.float64)
    P = np.zeros(mo.N_PARAMS)
    P[0:4] = F.mean(0)
    P[4:8] = 1.0 / F.std(0)
    P[8], P[9] = T.mean(), T.std()
    P[10:] = rng.normal(0, 0.3, mo.N_PARAMS - 10)
    loss, g = mo.loss_grad(P, F, T, bf16=False)
    assert np.all(g[:10] == 0)
    idx = np.concatenate([rng.choice(np.arange(10, mo.N_PARAMS), 250, replace=False), [mo.N_PARAMS - 1]])
    h = 1e-6
    fd = np.empty(len(idx))
    for n, i in enumerate(idx):
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd[n] = (mo.loss_grad(Pp, F, T, bf16=False)[0] - mo.loss_grad(Pm, F, T, bf16=False)[0]) / (2 * h)
    err = np.abs(fd - g[idx])
    assert np.max(err / (np.abs(fd) + 1e-6)) < 1e-4, (np.max(err), idx[np.argmax(err)])


def test_exact_solve_matches_brentq():
    optimize = pytest.importorskip("scipy.optimize")
    F, _ = mo.synthetic_samples(200, seed=4)
    F = F.astype(np.float64)
    e = mo.feature_s(F[:, 0]) - F[:, 1]
    x = mo.solve_x(e, F[:, 1])
    for n in range(len(F)):
        f0, f1 = e[n], F[n, 1]
        g = lambda v: 0.5 * (np.exp(f0 + v) - np.exp(f0 - v)) + v - f1  # noqa: E731
        with np.errstate(over="ignore"):  # g(f1) may be +inf: still a valid bracket end
            r = optimize.brentq(g, 0.0, f1, xtol=1e-14, rtol=1e-15, maxiter=500)
        assert abs(x[n] - r) <= 1e-11 * max(1.0, r)
    # log10 V from the closed form of the trial: V = coef e^e sinh x
    t = mo.target_log10V(F)
    V = 10 ** F[:, 2] * np.exp(e) * np.sinh(x)
    np.testing.assert_allclose(t, np.log10(V), rtol=0, atol=1e-9)
