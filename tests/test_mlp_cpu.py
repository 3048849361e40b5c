"""Pins of the numpy restatement of the MLP proposal model (oracle/mlp_ref.py)
that the GPU forward pass and gradients are checked against: analytic
gradients vs central finite differences in float64, and the mixture density
normalises to one.  CPU only."""
import numpy as np

import mlp_ref as R


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.3, size=R.NPARAMS)
    d = R.unpack(p)
    for c in range(R.NCOMP):  # sensible widths so that every component contributes
        d["b4"][6 * c + 3] = np.log(0.3)
        d["b4"][6 * c + 4] = np.log(0.4)
    return R.pack(d)


def test_mixture_nll_gradient_finite_differences():
    rng = np.random.default_rng(1)
    Y = rng.normal(scale=0.5, size=(5, R.NOUT))
    Y[:, 3::6] = rng.uniform(-2, 0, size=(5, R.NCOMP))
    Y[:, 4::6] = rng.uniform(-2, 0, size=(5, R.NCOMP))
    w, k = rng.uniform(size=5), rng.uniform(size=5)
    nll, dY = R.mix_nll(Y, w, k)
    h = 1e-6
    for i in range(5):
        for j in range(R.NOUT):
            Yp, Ym = Y.copy(), Y.copy()
            Yp[i, j] += h
            Ym[i, j] -= h
            fd = (R.mix_nll(Yp, w, k)[0][i] - R.mix_nll(Ym, w, k)[0][i]) / (2 * h)
            assert abs(fd - dY[i, j]) < 1e-6 * max(1.0, abs(fd)), (i, j, fd, dY[i, j])


def test_mixture_density_normalised():
    # a one-row output: integrate exp(-nll) over a wide (w, kappa) grid
    rng = np.random.default_rng(2)
    y = rng.normal(scale=0.3, size=(1, R.NOUT))
    y[0, 3::6] = np.log(0.1)
    y[0, 4::6] = np.log(0.15)
    g = np.linspace(-2, 3, 801)
    W, K = np.meshgrid(g, g, indexing="ij")
    nll, _ = R.mix_nll(np.repeat(y, W.size, axis=0), W.ravel(), K.ravel())
    assert abs(np.exp(-nll).sum() * (g[1] - g[0]) ** 2 - 1.0) < 1e-6


def test_backprop_finite_differences():
    p = _random_params(3)
    rng = np.random.default_rng(4)
    E = np.exp(rng.uniform(np.log(1e-3), np.log(1.0), size=7))
    X = R.features(E, 1e-3, 1.0)
    w, k = rng.uniform(size=7), rng.uniform(size=7)
    loss, g = R.loss_grad(p, X, w, k)
    h = 1e-6
    for idx in rng.choice(R.NPARAMS, 300, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (R.loss_grad(pp, X, w, k)[0] - R.loss_grad(pm, X, w, k)[0]) / (2 * h)
        assert abs(fd - g[idx]) < 1e-6 * max(1.0, abs(fd)), (idx, fd, g[idx])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159265, 65504.0, 1e-40])
    r = R.to_bf16(x)
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == 1.0078125
    assert r[4] == -3.140625 and r[5] == 65536.0
