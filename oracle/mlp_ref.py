"""TEST INFRASTRUCTURE ONLY: numpy restatement of the project's MLP proposal
model (ncrystal_amd/csrc/nc_mlp.hip), used by tests/ to check the GPU
forward pass and gradients.  The model has no counterpart in NCrystal.

Model: 8 energy features -> 3 hidden layers of 64 (SiLU) -> 48 outputs =
8 bivariate Gaussian components over (w, kappa): per component logit, mean_w,
mean_kappa, log sigma_w, log sigma_kappa (clamped to [-9, 3]) and c with
correlation rho = 0.95 tanh(c).  Loss: mean negative log-likelihood.

`bf16=True` rounds every matrix-product operand to bfloat16 (round to
nearest even) as the matrix-core kernels do; sums are kept in float64.
"""
import numpy as np

NIN, NH, NCOMP = 8, 64, 8
NOUT = 6 * NCOMP
OFF = {}
_o = 0
for _name, _size in (("W1", NIN * NH), ("b1", NH), ("W2", NH * NH), ("b2", NH), ("W3", NH * NH), ("b3", NH),
                     ("W4", NH * NOUT), ("b4", NOUT)):
    OFF[_name] = (_o, _size)
    _o += _size
NPARAMS = _o


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    g = lambda k: p[OFF[k][0]:OFF[k][0] + OFF[k][1]]
    return dict(W1=g("W1").reshape(NIN, NH), b1=g("b1"), W2=g("W2").reshape(NH, NH), b2=g("b2"),
                W3=g("W3").reshape(NH, NH), b3=g("b3"), W4=g("W4").reshape(NH, NOUT), b4=g("b4"))


def pack(d):
    out = np.zeros(NPARAMS)
    for k, (o, s) in OFF.items():
        out[o:o + s] = np.asarray(d[k]).ravel()
    return out


def to_bf16(x):
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    v = (((u + r) >> 16) << 16).astype(np.uint32)
    return v.view(np.float32).astype(np.float64)


def features(E, emin, emax):
    x = (np.log(np.asarray(E, dtype=np.float64)) - np.log(emin)) / (np.log(emax) - np.log(emin))
    pi = np.pi
    return np.stack([x, np.sin(pi * x), np.cos(pi * x), np.sin(2 * pi * x), np.cos(2 * pi * x),
                     np.sin(3 * pi * x), np.cos(3 * pi * x), x * x], axis=1)


def _mm(a, b, bf16):
    return (to_bf16(a) @ to_bf16(b)) if bf16 else a @ b


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, bf16=False):
    d = unpack(p)
    Z1 = _mm(X, d["W1"], bf16) + d["b1"]
    H1 = silu(Z1)
    Z2 = _mm(H1, d["W2"], bf16) + d["b2"]
    H2 = silu(Z2)
    Z3 = _mm(H2, d["W3"], bf16) + d["b3"]
    H3 = silu(Z3)
    Y = _mm(H3, d["W4"], bf16) + d["b4"]
    return Y, (X, Z1, H1, Z2, H2, Z3, H3)


def mix_nll(Y, w, kap):
    """per-row negative log-likelihood and its gradient with respect to Y"""
    n = Y.shape[0]
    Yc = Y.reshape(n, NCOMP, 6)
    logit = Yc[:, :, 0]
    lse = np.log(np.sum(np.exp(logit - logit.max(1, keepdims=True)), 1)) + logit.max(1)
    lsw = np.clip(Yc[:, :, 3], -9, 3)
    lsk = np.clip(Yc[:, :, 4], -9, 3)
    sw, sk = np.exp(lsw), np.exp(lsk)
    th = np.tanh(Yc[:, :, 5])
    rho = 0.95 * th
    om = 1 - rho * rho
    zw = (w[:, None] - Yc[:, :, 1]) / sw
    zk = (kap[:, None] - Yc[:, :, 2]) / sk
    q = (zw * zw - 2 * rho * zw * zk + zk * zk) / om
    lp = (logit - lse[:, None]) - np.log(2 * np.pi) - lsw - lsk - 0.5 * np.log(om) - 0.5 * q
    ll = np.log(np.sum(np.exp(lp - lp.max(1, keepdims=True)), 1)) + lp.max(1)
    g = np.exp(lp - ll[:, None])
    pik = np.exp(logit - lse[:, None])
    a = (zw - rho * zk) / om
    b = (zk - rho * zw) / om
    dY = np.zeros_like(Yc)
    dY[:, :, 0] = pik - g
    dY[:, :, 1] = -g * a / sw
    dY[:, :, 2] = -g * b / sk
    dY[:, :, 3] = np.where((Yc[:, :, 3] < -9) | (Yc[:, :, 3] > 3), 0.0, -g * (a * zw - 1))
    dY[:, :, 4] = np.where((Yc[:, :, 4] < -9) | (Yc[:, :, 4] > 3), 0.0, -g * (b * zk - 1))
    dlr = (rho + zw * zk - rho * q) / om
    dY[:, :, 5] = -g * dlr * 0.95 * (1 - th * th)
    return -ll, dY.reshape(n, NOUT)


def loss_grad(p, X, w, kap, bf16=False):
    """mean loss and gradient with respect to the packed parameters"""
    n = X.shape[0]
    d = unpack(p)
    Y, (X, Z1, H1, Z2, H2, Z3, H3) = forward(p, X, bf16)
    nll, dY = mix_nll(Y, w, kap)
    dY = dY / n
    gW4 = _mm(H3.T, dY, bf16)
    gb4 = dY.sum(0)
    dA = _mm(dY, d["W4"].T, bf16) * silu_grad(Z3)
    gW3 = _mm(H2.T, dA, bf16)
    gb3 = dA.sum(0)
    dB = _mm(dA, d["W3"].T, bf16) * silu_grad(Z2)
    gW2 = _mm(H1.T, dB, bf16)
    gb2 = dB.sum(0)
    dA = _mm(dB, d["W2"].T, bf16) * silu_grad(Z1)
    gW1 = _mm(X.T, dA, bf16)
    gb1 = dA.sum(0)
    return nll.mean(), pack(dict(W1=gW1, b1=gb1, W2=gW2, b2=gb2, W3=gW3, b3=gb3, W4=gW4, b4=gb4))


def mixture_params(y):
    """component weights, means, sigmas and correlations of one output row"""
    Yc = np.asarray(y, dtype=np.float64).reshape(NCOMP, 6)
    logit = Yc[:, 0]
    pi = np.exp(logit - logit.max())
    pi /= pi.sum()
    return dict(pi=pi, mu=Yc[:, 1:3], sig=np.exp(np.clip(Yc[:, 3:5], -9, 3)), rho=0.95 * np.tanh(Yc[:, 5]))
