"""Learned Newton guess on the GPU, through the C ABI: the forward kernels
(matrix-core wavefront form and scalar form), the gradient kernel and the
training loop against the FP64 oracle (oracle/mlp_oracle.py), and the
rate-and-state solve started from the trained guess against plain Newton
and against the oracle."""
import json

import numpy as np
import pytest

import cases
from oracle import mlp_oracle as mo
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

dyn = pytest.importorskip("dynfd3dcuda_amd")
from dynfd3dcuda_amd import mlp  # noqa: E402


def random_params(F, T, seed):
    """Initialised network with non-zero biases, so that both ReLU sides and
    all weights are exercised."""
    rng = np.random.default_rng(seed)
    P = mlp.init_params(F, T, seed).astype(np.float64)
    P[mo.OFF["b1"]:mo.OFF["W2"]] = rng.normal(0, 0.3, 64)
    P[mo.OFF["b2"]:mo.OFF["w3"]] = rng.normal(0, 0.3, 64)
    P[mo.OFF["b3"]] = 0.1
    return P.astype(np.float32)


@pytest.mark.parametrize("n", [1, 63, 1000, 4097])
def test_forward_matches_oracle(n):
    """Both GPU forms against the oracle with the GPU's roundings (W2 and
    the first hidden layer in bf16, FP32 elsewhere): the remaining
    differences are FP32 sums against FP64 sums and the bf16 rounding of a
    first-layer value that falls on the other side of a rounding boundary;
    bound: 2e-4 of the output scale (norm-wise) and 2e-3 (any sample)."""
    Fa, Ta = mo.synthetic_samples(4097, seed=11)
    P = random_params(Fa, Ta, 5)  # standardisation from the whole set
    F = Fa[:n]
    ref = mo.forward(P, F, bf16=True)
    scale = float(P[9])
    for mode in (0, 1):
        out = mlp.forward(P, F, mode).astype(np.float64)
        err = np.abs(out - ref) / scale
        assert np.sqrt(np.mean(err ** 2)) < 2e-4 and err.max() < 2e-3, (mode, err.max())


def test_forward_known_network_layout():
    """A network with exactly representable weights computing a known,
    operand-orientation-sensitive function (the oracle pins it): the
    matrix-core form must match it to FP32 rounding, which a wrong lane
    layout of any operand would break."""
    from test_mlp_oracle import _known_network
    P, _ = _known_network()
    rng = np.random.default_rng(3)
    F = rng.normal(0, 3, (640, 4)).astype(np.float32)
    ref = mo.forward(P, F, bf16=True)
    for mode in (0, 1):
        out = mlp.forward(P.astype(np.float32), F, mode)
        np.testing.assert_allclose(out, ref, rtol=0, atol=1e-4)


def test_gradient_matches_oracle():
    """GPU gradient (FP32 accumulation over 64-sample chunks and atomics)
    against the oracle's FP64 gradient with the same forward roundings:
    norm-wise 1e-4 relative; loss to 1e-5 relative."""
    F, T = mo.synthetic_samples(3001, seed=12)
    P = random_params(F, T, 6)
    loss, g = mlp.grad(P, F, T)
    lref, gref = mo.loss_grad(P, F, T, bf16=True)
    assert abs(loss - lref) <= 1e-5 * lref, (loss, lref)
    assert np.all(g[:10] == 0)
    e = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    assert e < 1e-4, e
    for k in ("W1", "b1", "W2", "b2", "w3"):
        a, b = mo.unpack(g)[k], mo.unpack(gref)[k]
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b) + 1e-12, k


def test_training_fits_exact_solve():
    """Train on the GPU on exact solutions of the rate-state equation
    (oracle) and check the fit on held-out samples, in log10 slip rate and
    in the Newton unknown x = friction / a that the guess starts from."""
    F, T = mo.synthetic_samples(200_000, seed=13)
    P, hist = mlp.train(F, T, epochs=12, batch=4096, lr=3e-3, seed=1)
    Fv, Tv = mo.synthetic_samples(20_000, seed=14)
    pred = mlp.forward(P, Fv, 0).astype(np.float64)
    err = np.abs(pred - Tv)
    # x from log10 V: V = coef e^e sinh x with e = s - f1
    e = mo.feature_s(Fv[:, 0]) - Fv[:, 1]
    x_true = mo.solve_x(e, Fv[:, 1].astype(np.float64))
    x_pred = np.arcsinh(10.0 ** (pred - Fv[:, 2] - e / np.log(10.0)))
    report = {"loss_first": float(hist[0]), "loss_last": float(hist[-1]),
              "log10V_median_err": float(np.median(err)), "log10V_p99_err": float(np.quantile(err, 0.99)),
              "x_median_err": float(np.median(np.abs(x_pred - x_true)))}
    print(json.dumps(report))
    assert np.all(np.isfinite(hist))
    assert hist[-1] < hist[0] / 20, report
    assert report["log10V_median_err"] < 0.2, report


def test_rs_solve_with_learned_guess():
    """Rate-state rupture: collect samples with plain Newton (NewtonMode 1),
    train on the GPU, rerun the same steps with the learned guess
    (NewtonMode 2).  The solve stops on the same residual rule, so the
    outputs agree with plain Newton and with the oracle to the tolerances
    of the plain rate-state test, while the Newton iterations per solve are
    reported for both."""
    nt = 30
    base = cases.rs_case(NewtonMode=1)
    s1 = dyn.Solver(base["params"])
    cases.apply_to_solver(base, s1)
    s1.setup()
    F, T = mlp.collect(s1, 0, nt, every=1)
    st1 = s1.newton_stats()
    P, hist = mlp.train(F, T, epochs=40, batch=2048, lr=3e-3, seed=2)
    case = cases.rs_case(NewtonMode=2)
    s2 = dyn.Solver(case["params"])
    cases.apply_to_solver(case, s2)
    s2.set_newton_mlp(P)
    s2.setup()
    for it in range(nt):
        s2.step(it)
    st2 = s2.newton_stats()
    o = Oracle(base["params"])
    cases.apply_to_oracle(base, o)
    o.setup()
    for it in range(nt):
        o.step(it)
    rep = {"samples": len(T), "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
           "plain_iters_per_solve": st1["iterations"] / st1["solves"],
           "mlp_iters_per_solve": st2["iterations"] / st2["solves"],
           "plain_fallbacks": st1["fallbacks"], "mlp_fallbacks": st2["fallbacks"]}
    for name in ("Vs1", "slip", "State", "Ts1"):
        a, b = s2.get(name).astype(np.float64), s1.get(name).astype(np.float64)
        rep[f"{name}_vs_plain"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        c = o.get(name).astype(np.float64)
        rep[f"{name}_vs_oracle"] = float(np.linalg.norm(a - c) / np.linalg.norm(c))
    print(json.dumps(rep))
    assert st1["solves"] == st2["solves"] > 0
    for name, tol in (("Vs1", 1e-3), ("Ts1", 1e-4), ("slip", 1e-3), ("State", 1e-4)):
        assert rep[f"{name}_vs_oracle"] < tol, (name, rep)
        assert rep[f"{name}_vs_plain"] < tol, (name, rep)


def _newton_iterations(x0, e, r, tol=1e-10, cap=100):
    """Iterations of Newton on g(x) = e^e sinh x + x - r (the rate-state
    solve divided by coef) from x0 until |g| <= tol * max(r, x): the
    residual rule of NewtonMode 1 and 2."""
    x = np.array(x0, np.float64)
    it = np.zeros(len(x), np.int64)
    done = np.zeros(len(x), bool)
    for _ in range(cap):
        ep, em = np.exp(np.minimum(e + x, 700.0)), np.exp(e - x)
        g = 0.5 * (ep - em) + x - r
        done |= np.abs(g) <= tol * np.maximum(np.abs(r), np.abs(x))
        if done.all():
            break
        x = np.where(done, x, x - g / (0.5 * (ep + em) + 1.0))
        it += ~done
    return it


def test_learned_guess_cuts_iterations_where_plain_guess_is_poor():
    """Where the previous slip rate is a poor predictor of the new one (the
    synthetic set draws it 0.5 decades off, as at a rupture front), Newton
    started from the trained network's guess (the better of it and the plain
    guess, as the solver chooses) needs fewer iterations, on average and in
    the worst case, than from the plain guess asinh(V_prev / (2 V0 e^(-psi/a))).  Network trained on the GPU;
    iterations counted in FP64 on the host with the solver's residual rule."""
    F, T = mo.synthetic_samples(200_000, seed=15)
    P, _ = mlp.train(F, T, epochs=12, batch=4096, lr=3e-3, seed=3)
    Fv, _ = mo.synthetic_samples(20_000, seed=16)
    lv = mlp.forward(P, Fv, 0).astype(np.float64)
    e = mo.feature_s(Fv[:, 0]) - Fv[:, 1]
    r = Fv[:, 1].astype(np.float64)
    ln10 = np.log(10.0)

    def x_of(log10v):  # x with coef e^e sinh x = V, V = 10^log10v
        y = (log10v - Fv[:, 2]) * ln10 - e  # ln(V / (coef e^e))
        return np.where(y > 20, y + np.log(2.0), np.arcsinh(np.exp(np.minimum(y, 700.0))))

    def resid(x):
        return np.abs(0.5 * (np.exp(np.minimum(e + x, 700.0)) - np.exp(e - x)) + x - r)

    x_plain, x_net = x_of(Fv[:, 3].astype(np.float64)), x_of(lv)
    x_best = np.where(resid(x_net) < resid(x_plain), x_net, x_plain)
    it_plain = _newton_iterations(x_plain, e, r)
    it_best = _newton_iterations(x_best, e, r)
    rep = {"plain_mean": float(it_plain.mean()), "learned_mean": float(it_best.mean()),
           "plain_max": int(it_plain.max()), "learned_max": int(it_best.max()),
           "plain_p99": float(np.quantile(it_plain, 0.99)), "learned_p99": float(np.quantile(it_best, 0.99))}
    print(json.dumps(rep))
    # locked nodes (the solve nearly linear) converge in one step from any
    # guess; the gain is at sliding nodes, where the plain guess can take tens
    # of steps: the worst case is what sets a wavefront's time
    assert rep["learned_mean"] < rep["plain_mean"], rep
    assert rep["learned_max"] <= rep["plain_max"] // 2, rep
