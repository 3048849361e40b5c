"""The MLP proposal model on the GPU: the BF16 matrix-core product (operand
lane layout) is exact on BF16-representable integers, the forward pass and
the gradients of the mixture likelihood agree with the numpy restatement
(which rounds the same operands to BF16), and training on samples of the
plain S(alpha,beta) sampler lowers the loss."""
import numpy as np
import pytest

import mlp_ref as R
from ncrystal_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return core


@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (37, 23, 70), (64, 48, 64), (5, 64, 8), (64, 64, 3000)])
def test_gemm_layout_exact(core, M, N, K):
    rng = np.random.default_rng(M * 1000 + N + K)
    # small integers: exact in BF16, products and sums exact in FP32
    A = rng.integers(-8, 9, size=(M, K)).astype(np.float32)
    B = rng.integers(-8, 9, size=(K, N)).astype(np.float32)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert np.array_equal(core.debug_gemm_bf16(A, B), ref)
    # transposed operands and accumulation into an existing C
    C0 = rng.integers(-4, 5, size=(M, N)).astype(np.float32)
    assert np.array_equal(core.debug_gemm_bf16(A.T.copy(), B, ta=True), ref)
    assert np.array_equal(core.debug_gemm_bf16(A, B.T.copy(), tb=True), ref)
    assert np.array_equal(core.debug_gemm_bf16(A, B, C0, accum=True), ref + C0)


def test_forward_and_gradient_vs_numpy(core):
    m = core.MLPProposal(1e-3, 1.0, seed=3)
    p = m.params().astype(np.float64)
    rng = np.random.default_rng(5)
    n = 2048
    E = np.exp(rng.uniform(np.log(1e-3), np.log(1.0), size=n))
    X = R.features(E, 1e-3, 1.0)
    Yg = m.forward(E).astype(np.float64)
    Yr, _ = R.forward(p, X, bf16=True)
    scale = np.abs(Yr).max()
    print("forward max abs diff / scale", np.abs(Yg - Yr).max() / scale)
    assert np.abs(Yg - Yr).max() <= 1e-2 * scale
    w, k = rng.uniform(size=n), rng.uniform(size=n)
    lg, gg = m.loss_grad(E, w, k)
    lr, gr = R.loss_grad(p, X, w, k, bf16=True)
    print("loss gpu", lg, "numpy", lr)
    assert abs(lg - lr) <= 1e-3 * abs(lr)
    gg = gg.astype(np.float64)
    for name, (o, s) in R.OFF.items():
        sl = slice(o, o + s)
        err = np.abs(gg[sl] - gr[sl]).max() / (np.abs(gr[sl]).max() + 1e-12)
        print("grad", name, "max rel diff", err)
        assert err <= 2e-2, name
    # the numpy reference without BF16 agrees to the BF16 level: the rounding
    # is the only difference between the two restatements
    _, g64 = R.loss_grad(p, X, w, k, bf16=False)
    assert np.abs(g64 - gr).max() <= 0.05 * np.abs(gr).max()


def test_training_lowers_loss(core):
    d = synth.sab_table_c3()
    sab = core.create_sabscatter(d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    m = core.MLPProposal(1e-3, 1.0, seed=1)
    first = m.train_sab(sab, steps=20, batch=4096, lr=3e-3, seed=2)
    last = m.train_sab(sab, steps=600, batch=8192, lr=3e-3, seed=3)
    print("mean loss: first steps", first, "after training", last)
    assert np.isfinite(last) and last < first - 0.3
