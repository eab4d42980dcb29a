"""Scattering surrogate (MLP 2 -> 64 -> 64 -> 1, tanh): the FP64 numpy
restatement pinned by finite differences, the device gradient against it, the
trained model against the MAGIC scattering of the oracle, and an ablation of
FP32+surrogate cascades against FP32 and FP64 ones."""
from __future__ import annotations

import math

import numpy as np
import pytest

import mlp_ref as M
import oracle_lib as O
from mytrim_amd import configs, gpu

EPS_MIN, EPS_MAX, B_MAX = 1e-7, 10.0, 40.0


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = gpu.Engine.surrogate_init(seed).astype(np.float64) * scale
    p[M.OFF_B1:M.OFF_W2] = rng.uniform(-0.5, 0.5, 64)
    p[M.OFF_B2:M.OFF_W3] = rng.uniform(-0.5, 0.5, 64)
    p[M.OFF_B3] = 0.3
    return p


def magic_samples(n, seed):
    rng = np.random.default_rng(seed)
    eps = 10.0 ** rng.uniform(math.log10(EPS_MIN), math.log10(EPS_MAX), n)
    b = rng.uniform(0.0, B_MAX, n)
    ls2 = np.array([math.log(max(O.scatter(e, bb, 0)[0], 1e-300)) for e, bb in zip(eps, b)])
    return eps, b, ls2


# ----------------------------------------------------------------- CPU checks
def test_parameter_layout():
    assert M.N_PARAMS == gpu.N_SURROGATE_PARAMS == 4417
    p = gpu.Engine.surrogate_init(1)
    assert p.dtype == np.float32 and p.shape == (4417,)
    assert np.all(p[M.OFF_B1:M.OFF_W2] == 0) and p[M.OFF_B3] == 0
    assert np.abs(p[M.OFF_W2:M.OFF_B2]).max() <= math.sqrt(6 / 128)


def test_backprop_matches_finite_differences():
    """The numpy gradient (checker of the device kernels) against central
    differences of the numpy loss, on every parameter block."""
    rng = np.random.default_rng(5)
    p = random_params(5, scale=1.5)
    X = rng.uniform(-1, 1, (24, 2))
    T = rng.uniform(-1.5, 0.5, 24)
    _, g = M.loss_grad(p, X, T)
    idx = np.concatenate([np.arange(0, 128, 7), np.arange(128, 192, 5), np.arange(192, 4288, 97),
                          np.arange(4288, 4352, 9), np.arange(4352, 4417, 4), [4416]])
    h = 1e-6
    for q in idx:
        pp, pm = p.copy(), p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (M.loss_grad(pp, X, T)[0] - M.loss_grad(pm, X, T)[0]) / (2 * h)
        assert fd == pytest.approx(g[q], rel=1e-5, abs=1e-9), q


def test_numpy_forward_special_cases():
    """W3 = 0 gives y = b3; a single active path reduces to nested tanh."""
    p = np.zeros(M.N_PARAMS)
    p[M.OFF_B3] = 0.7
    X = np.array([[0.3, -0.2], [1.0, 1.0]])
    assert np.allclose(M.forward(p, X), 0.7)
    p[M.OFF_W1 + 0] = 2.0          # W1[0][0]
    p[M.OFF_W2 + 0] = -1.5         # W2[0][0]
    p[M.OFF_W3 + 0] = 0.5
    want = 0.7 + 0.5 * np.tanh(-1.5 * np.tanh(2.0 * X[:, 0]))
    assert np.allclose(M.forward(p, X), want, rtol=1e-15)


# ----------------------------------------------------------------- GPU checks
@pytest.fixture(scope="module")
def eng():
    e = gpu.Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="module")
def trained(eng):
    opts = gpu.Engine.surrogate_opts(EPS_MIN, EPS_MAX, B_MAX, n_train=1 << 22, epochs=10,
                                     batch=4096, lr=2e-3, seed=11)
    p, loss = eng.surrogate_train(opts, gpu.Engine.surrogate_init(3))
    print("surrogate epoch loss:", " ".join("%.3g" % x for x in loss))
    return opts, p, loss


@pytest.mark.gpu
def test_device_gradient_matches_numpy(eng):
    opts = gpu.Engine.surrogate_opts(EPS_MIN, EPS_MAX, B_MAX)
    p = random_params(7).astype(np.float32)
    eps, b, ls2 = magic_samples(777, 3)
    g, loss = eng.surrogate_grad(opts, p, eps, b, ls2)
    X = M.normalise(eps, b, EPS_MIN, EPS_MAX, B_MAX)
    want_loss, want = M.loss_grad(p.astype(np.float64), X, M.target(ls2))
    err = np.linalg.norm(g - want) / np.linalg.norm(want)
    print("gradient rel. error %.2e, loss %.6g vs %.6g" % (err, loss, want_loss))
    assert loss == pytest.approx(want_loss, rel=1e-5)
    assert err < 1e-4
    for lo, hi in [(0, 128), (128, 192), (192, 4288), (4288, 4352), (4352, 4416)]:
        blk = np.linalg.norm(g[lo:hi] - want[lo:hi]) / np.linalg.norm(want[lo:hi])
        assert blk < 1e-4, (lo, hi, blk)


@pytest.mark.gpu
def test_training_reduces_loss_and_fits_magic(eng, trained):
    opts, p, loss = trained
    assert np.all(np.isfinite(loss))
    assert loss[-1] < loss[0] / 10
    eps, b, ls2 = magic_samples(4000, 9)
    s2n = M.s2_of(p, eps, b, EPS_MIN, EPS_MAX, B_MAX)
    dl = np.abs(np.log(s2n) - ls2)
    print("trained surrogate |d ln s2|: median %.3g, p90 %.3g, max %.3g" %
          (np.median(dl), np.quantile(dl, 0.9), dl.max()))
    assert np.median(dl) < 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("bf16", [False, True])
def test_device_inference_matches_numpy(eng, trained, bf16):
    opts, p, _ = trained
    opts.bf16 = int(bf16)
    eng.surrogate_install(opts, p)
    eps, b, _ = magic_samples(3000, 13)
    got = eng.surrogate_eval(eps, b)
    want = M.s2_of(p.astype(np.float64), eps, b, EPS_MIN, EPS_MAX, B_MAX)
    d = np.abs(np.log(got) - np.log(want))
    print("bf16=%d device vs numpy |d ln s2|: median %.2e max %.2e" % (bf16, np.median(d), d.max()))
    if bf16:
        assert np.median(d) < 0.02 and d.max() < 0.5
    else:
        assert d.max() < 2e-3
    # outside the domain the device reports -1 (the MAGIC solve runs instead)
    out = eng.surrogate_eval(np.array([EPS_MAX * 2, 1e-9, 1.0]), np.array([1.0, 1.0, B_MAX * 1.5]))
    assert np.all(out == -1.0)


@pytest.mark.gpu
def test_surrogate_cascade_ablation(eng, trained):
    """C -> W at 2 MeV: FP32 + surrogate against FP32 MAGIC and FP64 MAGIC on
    the same histories (vacancies/ion, electronic share, mean range)."""
    opts, p, _ = trained
    opts.bf16 = 1
    eng.surrogate_install(opts, p)
    prob = configs.c5_c_w(1000, seed=43)
    res = {}
    for name, ar in [("fp64", gpu.MT_FP64), ("fp32", gpu.MT_FP32), ("surr", gpu.MT_FP32_SURROGATE)]:
        res[name] = eng.run(prob, per_history=True, arith=ar)
    n = prob.n_hist
    a = res["fp64"]
    for name, r in res.items():
        print("%s: %.1f ms, %.4g coll, surrogate %.4g, vac/ion %.2f, Eel/E0 %.4f, range %.1f" % (
            name, r["kernel_ms"], r["n_coll"], r["surrogate"], r["vac"] / n, r["Eel"] / r["E0"],
            r["range_sum"] / r["range_n"]))
    s = res["surr"]
    assert 0 < s["surrogate"] < s["n_coll"]
    assert res["fp64"]["surrogate"] == 0 and res["fp32"]["surrogate"] == 0
    sig = math.sqrt((np.var(a["h_vac"]) + np.var(s["h_vac"])) / n)
    assert abs(s["vac"] - a["vac"]) / n <= max(0.05 * a["vac"] / n, 3 * sig)
    assert s["Eel"] / s["E0"] == pytest.approx(a["Eel"] / a["E0"], rel=0.03)
    tot = s["Eel"] + s["Enuc"] + s["Eunf"] + s["Elost"] + s["Edrop"] + s["Evac"]
    assert tot == pytest.approx(s["E0"], rel=1e-11)


@pytest.mark.gpu
def test_surrogate_requires_install_and_universal_potential(eng):
    e2 = gpu.Engine(0)
    try:
        with pytest.raises(RuntimeError):
            e2.run(configs.c5_c_w(4, seed=1), arith=gpu.MT_FP32_SURROGATE)
        with pytest.raises(RuntimeError):
            e2.surrogate_eval(np.array([1.0]), np.array([1.0]))
    finally:
        e2.close()
