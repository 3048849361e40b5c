"""Exact envelope sampler of S(alpha,beta) (box rejection in kinematic
coordinates) on the GPU, on the uniform grid and on the grid placed by the
learned model.  Its samples follow the interpolated table density exactly, so
they are checked against (1) the bounds themselves (every box's upper bound
dominates the target and its lower bound stays below it on a fine grid of
the box, evaluated here with numpy), (2) the exact beta marginal of the
interpolated table obtained by numerical integration, and (3) the plain
sampler of the CPU checker by two-sample KS.  The squeeze (acceptance below
the lower bound without evaluating the table) must not change any sample.
Neutrons the envelope sampler hands back fall back to the plain sampler on
their own Philox stream, so they must equal the plain GPU sampler exactly."""
import math

import numpy as np
import pytest
from scipy import stats

from ncrystal_amd import synth

pytestmark = pytest.mark.gpu

KB = 8.6173303e-5


@pytest.fixture(scope="module")
def env():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return _capi, core


def _table():
    d = synth.sab_table_c3()
    return d, d["alpha"], d["beta"], d["sab"].reshape(len(d["beta"]), len(d["alpha"]))


def _interp_s(a, b, S, al, be):
    """SABCellEval (alpha first: log-linear with linear fallback, then linear in beta)"""
    ia = np.clip(np.searchsorted(a, al, side="right") - 1, 0, len(a) - 2)
    ib = np.clip(np.searchsorted(b, be, side="right") - 1, 0, len(b) - 2)
    a0, a1 = a[ia], a[ia + 1]

    def row(j):
        s0, s1 = S[j, ia], S[j, ia + 1]
        with np.errstate(divide="ignore", invalid="ignore"):
            ll = np.exp(np.log(s0) + (np.log(s1) - np.log(s0)) * (al - a0) / (a1 - a0))
        lin = s0 + (s1 - s0) * (al - a0) / (a1 - a0)
        return np.where(s0 * s1 == 0.0, lin, ll)

    tb = (be - b[ib]) / (b[ib + 1] - b[ib])
    v = row(ib) + (row(ib + 1) - row(ib)) * tb
    inside = (al >= a[0]) & (al <= a[-1]) & (be >= b[0]) & (be <= b[-1])
    return np.where(inside, v, 0.0)


def _check_bounds(g, d, a, b, S, intervals, nsample=200, seed=1):
    """every sampled box: lower bound <= t <= upper bound on a 24 x 24 grid
    of the box; returns per interval the ratio target mass / envelope mass"""
    node, boxes = g.envelope_tables()
    kT = KB * d["temperature"]
    rng = np.random.default_rng(seed)
    eff = {}
    for n in intervals:
        eps = node[n + 1] / kT
        wmax = math.sqrt(b[-1] + eps)
        jf = 8 * math.sqrt(eps) * wmax ** 3
        bx = boxes[n]
        for j in rng.choice(len(bx), nsample, replace=False):
            w0, w1, k0, k1, hi, lo = bx[j]
            W, K = np.meshgrid(w0 + (w1 - w0) * np.linspace(0, 1, 24), k0 + (k1 - k0) * np.linspace(0, 1, 24),
                               indexing="ij")
            r = W * wmax
            t = _interp_s(a, b, S, r * r + eps - 2 * math.sqrt(eps) * r * (1 - 2 * K), r * r - eps) * jf * W * W
            assert np.all(t <= hi * (1 + 1e-12)), (n, j)
            assert np.all(t >= lo * (1 - 1e-12)), (n, j)
        # target mass on a fine midpoint grid of the unit square vs envelope mass
        W, K = np.meshgrid((np.arange(512) + 0.5) / 512, (np.arange(256) + 0.5) / 256, indexing="ij")
        r = W * wmax
        t = _interp_s(a, b, S, r * r + eps - 2 * math.sqrt(eps) * r * (1 - 2 * K), r * r - eps) * jf * W * W
        area = (bx[:, 1] - bx[:, 0]) * (bx[:, 3] - bx[:, 2])
        assert abs(area.sum() - 1.0) < 1e-6
        eff[n] = (t.mean() / np.sum(bx[:, 4] * area), np.sum(bx[:, 5] * area) / np.sum(bx[:, 4] * area))
    return eff


def test_envelope_bounds_target(env):
    C, core = env
    d, a, b, S = _table()
    g = core.create_sabscatter(d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    nn = 64
    g.use_envelope_sampler(nintervals=nn, emin=1e-3, emax=1.0, nbox_w=32, nbox_k=16)
    eff = _check_bounds(g, d, a, b, S, (0, 13, 40, nn - 1))
    for n, (acc, sq) in eff.items():
        # the bounds are not loose on average (coarse 32 x 16 boxes here)
        print("interval", n, "acceptance", acc, "squeeze fraction", sq)
        assert acc > 0.45 and sq > 0.2
    st = g.envelope_stats()
    assert np.all(st[:, 2] == 1.0) and np.array_equal(st[:, 0], st[:, 1])


def test_envelope_squeeze_changes_nothing(env):
    C, core = env
    d, a, b, S = _table()
    g = core.create_sabscatter(d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    g.use_envelope_sampler()
    e = synth.energies_logu(1000000, 1e-3, 1.0, seed=8)
    de = C.DeviceArray.from_numpy(e)
    st = C.Stream()
    out = []
    for on in (1, 0):
        g.set_envelope_squeeze(on)
        g.setRandomSeed(12)
        da, db = C.DeviceArray(e.size), C.DeviceArray(e.size)
        g.dev_sample_alphabeta_envelope(de.ptr, e.size, da.ptr, db.ptr, stream=st.ptr)
        out.append((da.numpy(), db.numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def _beta_cdf_exact(a, b, S, E, kT, sub=16):
    eps = E / kT
    af = np.concatenate([a[:-1, None] + (a[1:] - a[:-1])[:, None] * (np.arange(32) / 32)[None, :]]).ravel()
    af = np.append(af, a[-1])
    bq = np.concatenate([b[:-1, None] + (b[1:] - b[:-1])[:, None] * (np.arange(sub) / sub)[None, :]]).ravel()
    bq = np.append(bq, b[-1])
    bq = bq[bq > -eps]
    # marginal density in beta: integral over the kinematic alpha range of the interpolated S
    m = np.empty(len(bq))
    for j0 in range(0, len(bq), 2000):
        bb = bq[j0:j0 + 2000]
        lo = (np.sqrt(eps + bb) - math.sqrt(eps)) ** 2
        hi = (np.sqrt(eps + bb) + math.sqrt(eps)) ** 2
        AL = np.clip(af[None, :], lo[:, None], hi[:, None])
        val = _interp_s(a, b, S, AL, np.broadcast_to(bb[:, None], AL.shape))
        m[j0:j0 + 2000] = np.sum(0.5 * (val[:, 1:] + val[:, :-1]) * np.diff(AL, axis=1), axis=1)
    F = np.concatenate([[0.0], np.cumsum(0.5 * (m[1:] + m[:-1]) * np.diff(bq))])
    return bq, F / F[-1]


def test_envelope_sampler_distribution(env, oracle):
    C, core = env
    d, a, b, S = _table()
    args = (d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    g = core.create_sabscatter(*args)
    g.use_envelope_sampler()
    o = oracle.sab(*args)
    kT = KB * d["temperature"]
    st = C.Stream()
    n = 2000000
    for E in (2e-3, 0.025, 0.2, 0.9):
        e = np.full(n, E)
        de = C.DeviceArray.from_numpy(e)
        da, db = C.DeviceArray(n), C.DeviceArray(n)
        g.setRandomSeed(3)
        nfb = g.dev_sample_alphabeta_envelope(de.ptr, n, da.ptr, db.ptr, stream=st.ptr)
        al, be = da.numpy(), db.numpy()
        assert nfb < 1e-3 * n
        # kinematically valid and inside the table
        lo = (np.sqrt(E / kT + be) - math.sqrt(E / kT)) ** 2
        hi = (np.sqrt(E / kT + be) + math.sqrt(E / kT)) ** 2
        assert np.all(be > -E / kT) and np.all(al >= lo * (1 - 1e-9)) and np.all(al <= hi * (1 + 1e-9))
        # beta marginal vs the exact marginal of the interpolated table
        bq, F = _beta_cdf_exact(a, b, S, E, kT)
        bs = np.sort(be)
        ks_exact = np.max(np.abs(np.searchsorted(bs, bq, side="right") / n - F))
        # two-sample KS vs the plain sampler of the CPU checker
        ra, rb, nerr = o.sample_ab(e, seed=5, nthreads=8)
        ks_b = stats.ks_2samp(be, rb).statistic
        ks_a = stats.ks_2samp(al, ra).statistic
        print(f"E={E}: fallback {nfb}, KS(beta vs exact marginal) {ks_exact:.2e}, "
              f"KS vs plain: beta {ks_b:.2e} alpha {ks_a:.2e}")
        assert ks_exact < 1.6 / math.sqrt(n) + 2e-4
        assert ks_b < 2.5 / math.sqrt(n) + 3e-4 and ks_a < 2.5 / math.sqrt(n) + 3e-4


def test_envelope_sampler_iso_and_fallback(env):
    C, core = env
    d, a, b, S = _table()
    args = (d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    plain = core.create_sabscatter(*args)
    kbs = core.create_sabscatter(*args)
    # covered range [1e-3, 0.5) eV: neutrons outside take the plain sampler from their first draw
    kbs.use_envelope_sampler(nintervals=256, emin=1e-3, emax=0.5)
    e = synth.energies_logu(400000, 1e-4, 1.0, seed=4)
    plain.setRandomSeed(21)
    kbs.setRandomSeed(21)
    pe, pm = plain.sampleScatterIsotropic(e)
    ke, km = kbs.sampleScatterIsotropic(e)
    out = (e < 1e-3) | (e >= 0.5)
    assert np.array_equal(pe[out], ke[out]) and np.array_equal(pm[out], km[out])
    assert np.all(np.abs(km) <= 1.0) and np.all(ke >= 0.0)
    inr = ~out
    assert stats.ks_2samp(pm[inr], km[inr]).statistic < 0.01
    assert stats.ks_2samp(pe[inr], ke[inr]).statistic < 0.01
    kbs.use_plain_sampler()
    kbs.setRandomSeed(21)
    ke2, km2 = kbs.sampleScatterIsotropic(e)
    assert np.array_equal(ke2, pe) and np.array_equal(km2, pm)


@pytest.fixture(scope="module")
def trained(env):
    C, core = env
    d, a, b, S = _table()
    args = (d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    trainer = core.create_sabscatter(*args)
    model = core.MLPProposal(1e-3, 1.0, seed=1)
    loss = model.train_sab(trainer, steps=1500, batch=8192, lr=3e-3, seed=1)
    print("trained mean loss", loss)
    return model


def test_learned_sampler_distribution_and_gain(env, oracle, trained):
    C, core = env
    d, a, b, S = _table()
    args = (d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    g = core.create_sabscatter(*args)
    g.use_learned_sampler(trained, nintervals=256, emin=1e-3, emax=1.0)
    st_ = g.envelope_stats()
    mass, uni, lam, sq = st_.T
    gain = uni / mass
    print("envelope mass uniform grid / learned grid: median %.3f min %.3f max %.3f; uniform blend weights %s; "
          "squeeze fraction median %.3f" % (np.median(gain), gain.min(), gain.max(),
                                           np.unique(lam, return_counts=True), np.median(sq / mass)))
    # the learned grid is never worse than the uniform grid it is chosen
    # against, and it gains over the method without the network (measured
    # median 1.065 on 128 x 32 boxes)
    assert np.all(mass <= uni)
    assert np.median(gain) > 1.03
    eff = _check_bounds(g, d, a, b, S, (0, 100, 255), nsample=100)
    print("learned grid acceptance / squeeze fraction", eff)
    o = oracle.sab(*args)
    kT = KB * d["temperature"]
    st = C.Stream()
    n = 2000000
    for E in (2e-3, 0.025, 0.2, 0.9):
        e = np.full(n, E)
        de = C.DeviceArray.from_numpy(e)
        da, db = C.DeviceArray(n), C.DeviceArray(n)
        g.setRandomSeed(4)
        nfb = g.dev_sample_alphabeta_envelope(de.ptr, n, da.ptr, db.ptr, stream=st.ptr)
        al, be = da.numpy(), db.numpy()
        assert nfb < 1e-3 * n
        bq, F = _beta_cdf_exact(a, b, S, E, kT)
        ks_exact = np.max(np.abs(np.searchsorted(np.sort(be), bq, side="right") / n - F))
        ra, rb, nerr = o.sample_ab(e, seed=6, nthreads=8)
        ks_b = stats.ks_2samp(be, rb).statistic
        ks_a = stats.ks_2samp(al, ra).statistic
        print(f"learned E={E}: fallback {nfb}, KS(beta vs exact marginal) {ks_exact:.2e}, "
              f"KS vs plain: beta {ks_b:.2e} alpha {ks_a:.2e}")
        assert ks_exact < 1.6 / math.sqrt(n) + 2e-4
        assert ks_b < 2.5 / math.sqrt(n) + 3e-4 and ks_a < 2.5 / math.sqrt(n) + 3e-4
    # squeeze on and off give the same samples on the learned grid too
    e = synth.energies_logu(500000, 1e-3, 1.0, seed=9)
    de = C.DeviceArray.from_numpy(e)
    res = []
    for on in (1, 0):
        g.set_envelope_squeeze(on)
        g.setRandomSeed(13)
        da, db = C.DeviceArray(e.size), C.DeviceArray(e.size)
        g.dev_sample_alphabeta_envelope(de.ptr, e.size, da.ptr, db.ptr, stream=st.ptr)
        res.append((da.numpy(), db.numpy()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_learned_sampler_iso_fallback_identity(env, trained):
    C, core = env
    d, a, b, S = _table()
    args = (d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    plain = core.create_sabscatter(*args)
    lps = core.create_sabscatter(*args)
    lps.use_learned_sampler(trained, nintervals=128, emin=2e-3, emax=0.5)
    e = synth.energies_logu(400000, 1e-4, 1.0, seed=5)
    plain.setRandomSeed(22)
    lps.setRandomSeed(22)
    pe, pm = plain.sampleScatterIsotropic(e)
    le, lm = lps.sampleScatterIsotropic(e)
    out = (e < 2e-3) | (e >= 0.5)
    assert np.array_equal(pe[out], le[out]) and np.array_equal(pm[out], lm[out])
    inr = ~out
    assert stats.ks_2samp(pm[inr], lm[inr]).statistic < 0.01
    assert stats.ks_2samp(pe[inr], le[inr]).statistic < 0.01


def test_envelope_range_clipped_to_table_grid(env):
    """outside the table's energy grid the reference does not sample the
    table density (free-gas extender above, low-energy rule below): the
    envelope sampler's range is clipped to the grid and those neutrons get
    exactly the plain sampler's values"""
    C, core = env
    d, a, b, S = _table()
    args = (d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    plain = core.create_sabscatter(*args)
    kbs = core.create_sabscatter(*args)
    kbs.use_envelope_sampler(nintervals=64, emin=1e-9, emax=100.0)
    eg, _ = kbs.egrid()
    node, _ = kbs.envelope_tables()
    assert node[0] == eg[0] and node[-1] == eg[-1]
    e = np.concatenate([synth.energies_logu(100000, 1e-7, eg[0], seed=3),
                        synth.energies_logu(100000, eg[-1], 20.0, seed=4)])
    plain.setRandomSeed(5)
    kbs.setRandomSeed(5)
    pe, pm = plain.sampleScatterIsotropic(e)
    ke, km = kbs.sampleScatterIsotropic(e)
    assert np.array_equal(pe, ke) and np.array_equal(pm, km)
