"""Pins of the CPU checker (oracle/ncoracle.c) against the reference's own
golden data and against closed forms / brute force.  CPU only."""
import json
import math

import numpy as np
import pytest
from scipy import integrate, stats

from conftest import GOLDEN
from ncrystal_amd import synth

WL2EKIN = 0.081804209605330899
K2PI = 6.2831853071795864769252867665590057683943388


def test_philox_known_answers(oracle):
    # Random123 kat_vectors, philox4x32_10
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_u01_mapping(oracle):
    L = oracle.lib()
    assert L.orc_u64_to_fp01(0) == 1.0
    # largest integer gives the smallest value 2^-53 - 0x7FF*2^-64 (NCDefs.hh:1308-1329)
    assert L.orc_u64_to_fp01(2**64 - 1) == pytest.approx(2.0**-53 - 0x7FF * 2.0**-64, rel=1e-15)
    v = oracle.rng_fill(1, 0, 0, 200000)
    assert v.min() > 0.0 and v.max() <= 1.0
    assert stats.kstest(v, "uniform").pvalue > 1e-4
    # distinct neutrons / calls give distinct streams
    assert not np.array_equal(oracle.rng_fill(1, 0, 1, 8), oracle.rng_fill(1, 0, 2, 8))
    assert not np.array_equal(oracle.rng_fill(1, 1, 1, 8), oracle.rng_fill(1, 0, 1, 8))


def test_debye_msd_golden(oracle):
    g = json.loads((GOLDEN / "debye_msd.json").read_text())
    for v in g["values"]:
        got = math.sqrt(oracle.debye_msd(v["debye_temp"], v["temperature"], v["mass"]))
        assert got == pytest.approx(v["sqrt_msd"], rel=6e-6), v


def test_debye_msd_closed_form(oracle):
    # T=0: msd = 3 hbar^2 c^2 / (4 M kB T_D) exactly (shape function 1/4)
    hbar = 4.135667662e-15 / (2 * math.pi)
    c = 299792458e10
    for td, m in [(100.0, 1.00794), (410.0, 26.98), (2000.0, 238.0)]:
        exact = 3 * hbar * hbar * c * c / (931494095.17 * 8.6173303e-5) / (m * td) * 0.25
        assert oracle.debye_msd(td, 0.0, m) == pytest.approx(exact, rel=1e-14)
    # high-T limit: msd -> 3 hbar^2 T /(M kB T_D^2) * (1 + T_D^2/(36 T^2))
    td, m, t = 100.0, 26.98, 5000.0
    lim = 3 * hbar * hbar * c * c / (931494095.17 * 8.6173303e-5) / (m * td) * (t / td) * (1 + (td / t) ** 2 / 36)
    assert oracle.debye_msd(td, t, m) == pytest.approx(lim, rel=1e-6)


def test_alpha_integral_known_answers(oracle):
    g = json.loads((GOLDEN / "sab_alpha_integrals.json").read_text())
    L = oracle.lib()
    for y, ref in g["values"]:
        s1, s2 = (1.0 - y) / 2.0, (1.0 + y) / 2.0
        got = L.orc_integrate_alpha_interval_fast(0.0, s1, 1.0, s2, math.log(s1), math.log(s2))
        assert abs(got - ref) <= 0.5 * 1e-14 * (abs(got) + abs(ref)), (y, got, ref)


def test_sincos_2pix(oracle):
    import ctypes
    L = oracle.lib()
    s, c = ctypes.c_double(), ctypes.c_double()
    for x in np.linspace(-3.7, 5.3, 2001):
        L.orc_sincos_2pix(float(x), ctypes.byref(s), ctypes.byref(c))
        fr = float(x) - math.floor(float(x))  # exact
        assert abs(s.value - math.sin(2 * math.pi * fr)) < 1e-15
        assert abs(c.value - math.cos(2 * math.pi * fr)) < 1e-15


def _al_nosg(oracle):
    g = json.loads((GOLDEN / "nosg_al_hkl.json").read_text())
    m = g["material"]
    msd = oracle.debye_msd(m["debye_temp"], m["temperature"], m["mass"])
    pos = synth._norm_positions(m["positions"])
    res = oracle.hkl_orthogonal(m["a"], m["a"], m["a"], [(m["csl"], msd, pos)], m["dcutoff"])
    return g, msd, res


def test_hkl_nosg_golden(oracle):
    g, msd, (hkl, mult, dsp, fsq) = _al_nosg(oracle)
    assert msd == pytest.approx(g["material"]["msd_printed"], rel=1e-6)
    fams = g["families"]
    assert len(hkl) == len(fams)
    for i, f in enumerate(fams):
        assert list(hkl[i]) == f["hkl"]
        assert mult[i] == f["mult"] == 2 * len(f["members"])
        assert dsp[i] == pytest.approx(f["d"], rel=6e-6)
        assert fsq[i] == pytest.approx(f["fsq"], rel=6e-6)


def test_hkl_bruteforce_multiplicity(oracle):
    """Sum of multiplicities = number of non-zero (h,k,l) with d >= dcutoff
    (all |F|^2 above the cut for the fcc-allowed reflections); fcc extinction
    rule (h,k,l all even or all odd) by brute force."""
    g, msd, (hkl, mult, dsp, fsq) = _al_nosg(oracle)
    a, dcut = g["material"]["a"], g["material"]["dcutoff"]
    n = int(a / dcut) + 1
    cnt = 0
    for h in range(-n, n + 1):
        for k in range(-n, n + 1):
            for l in range(-n, n + 1):
                if (h, k, l) == (0, 0, 0):
                    continue
                if len({h % 2, k % 2, l % 2}) != 1:
                    continue
                d = a / math.sqrt(h * h + k * k + l * l)
                if d >= dcut:
                    cnt += 1
    assert mult.sum() == cnt


def test_fcc_c1_survey_values(oracle):
    """fcc a=4.05, M=26.98, T_D=410 K, T=293 K, dcutoff 0.5: values recorded
    from the reference in the survey (first family and PowderBragg xs at 4 Aa)."""
    cr = synth.fcc_powder()
    msd = oracle.debye_msd(cr.species[0].debye_temp, cr.temperature, cr.species[0].mass)
    hkl, mult, dsp, fsq = oracle.hkl_orthogonal(cr.a, cr.b, cr.c, [(cr.species[0].csl, msd, cr.species[0].positions)],
                                                cr.dcutoff)
    assert list(hkl[0]) == [1, 1, 1] and mult[0] == 8
    assert fsq[0] == 1.7719496096281011
    # the no-space-group path computes d = 2pi*(1/|k|); the survey value came
    # from the space-group path, d = 2pi/|k| (NCFillHKL.cc:676): 1 ulp apart
    ksq = 3 * (K2PI / 4.05) ** 2
    assert K2PI / math.sqrt(ksq) == 2.3382685902179845
    assert dsp[0] == K2PI * (1.0 / math.sqrt(ksq))
    pb = oracle.powderbragg(cr.volume * cr.n_atoms, dsp, fsq * mult)
    e4 = WL2EKIN / 16.0
    assert pb.xs([e4])[0] == pytest.approx(1.6308398365212686, rel=1e-15)


def test_powderbragg_closed_form(oracle):
    cr = synth.fcc_powder()
    msd = oracle.debye_msd(410.0, 293.0, 26.98)
    hkl, mult, dsp, fsq = oracle.hkl_orthogonal(cr.a, cr.b, cr.c, [(cr.species[0].csl, msd, cr.species[0].positions)],
                                                cr.dcutoff)
    v0n = cr.volume * cr.n_atoms
    pb = oracle.powderbragg(v0n, dsp, fsq * mult)
    wl = np.linspace(0.3, 6.0, 4001)
    e = WL2EKIN / wl**2
    xs = pb.xs(e)
    c = dsp * fsq * mult * 0.5 / v0n
    exact = np.array([(c[2 * dsp >= w]).sum() * w * w for w in wl])
    np.testing.assert_allclose(xs, exact, rtol=1e-13, atol=0)
    # above the largest 2d the cross section vanishes
    assert pb.xs([WL2EKIN / (2 * dsp[0] * 1.0001) ** 2])[0] == 0.0
    assert pb.xs([0.0, -1.0, np.inf, np.nan]).tolist() == [0.0, 0.0, 0.0, 0.0]


def test_powderbragg_mu(oracle):
    # single family: mu = 1 - 2 (lambda/2d)^2 exactly
    pb = oracle.powderbragg(10.0, [2.0], [3.0])
    wl = 3.1
    e = WL2EKIN / wl**2
    oe, om, nerr = pb.sample_iso(np.full(1000, e))
    assert nerr == 0 and np.all(oe == e)
    np.testing.assert_allclose(om, 1.0 - 2.0 * (wl / 4.0) ** 2, rtol=1e-14)
    # family probabilities proportional to d*F^2*m (chi-square)
    d = np.array([3.0, 2.0, 1.5])
    fm = np.array([1.0, 2.0, 4.0])
    pb = oracle.powderbragg(10.0, d, fm)
    e = WL2EKIN / 1.0**2
    _, om, _ = pb.sample_iso(np.full(200000, e), seed=7)
    mus = 1.0 - 2.0 * (1.0 / (2 * d)) ** 2
    counts = np.array([np.sum(np.isclose(om, m, rtol=0, atol=1e-12)) for m in mus])
    assert counts.sum() == 200000
    p = d * fm / np.sum(d * fm)
    assert stats.chisquare(counts, p * 200000).pvalue > 1e-4


def test_elinc_closed_form(oracle):
    msd = np.array([0.01, 0.02, 0.005])
    bixs = np.array([2.0, 0.3, 1.1])
    scale = np.array([0.5, 0.25, 0.25])
    ei = oracle.elinc(msd, bixs, scale)
    e = np.geomspace(1e-6, 10.0, 1000)
    k = 16 * math.pi**2 / WL2EKIN
    exact = np.zeros_like(e)
    for m, b, s in zip(msd, bixs, scale):
        t = k * m * e
        exact += b * s * (-np.expm1(-t) / t)
    np.testing.assert_allclose(ei.xs(e), exact, rtol=1e-9)  # Taylor branch: 10 digits at t=0.01


def test_elinc_mu_distribution(oracle):
    msd = 0.01
    ei = oracle.elinc([msd], [1.0], [1.0])
    for e in (0.0005, 0.05, 1.0):
        a = 8 * math.pi**2 / WL2EKIN * e * msd
        _, om, _ = ei.sample_iso(np.full(100000, e), seed=3)
        cdf = lambda x: np.expm1(a * (np.asarray(x) + 1)) / np.expm1(2 * a)
        assert stats.kstest(om, cdf).pvalue > 1e-4


def _sab_small(oracle):
    d = synth.small_sab_table()
    return d, oracle.sab(d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])


def _loglin(a0, s0, a1, s1, x):
    if s0 > 0 and s1 > 0:
        return s0 * (s1 / s0) ** ((x - a0) / (a1 - a0))
    return s0 + (s1 - s0) * (x - a0) / (a1 - a0)


def _bruteforce_beta_xs(alpha, srow, alow, aupp):
    """integral over [alow,aupp] of the log-linear-in-alpha interpolant of one
    beta row (zero outside the grid), by adaptive quadrature per cell."""
    lo, hi = max(alow, alpha[0]), min(aupp, alpha[-1])
    if not hi > lo:
        return 0.0
    tot = 0.0
    for i in range(len(alpha) - 1):
        a0, a1 = alpha[i], alpha[i + 1]
        x0, x1 = max(lo, a0), min(hi, a1)
        if x1 > x0:
            tot += integrate.quad(lambda x: _loglin(a0, srow[i], a1, srow[i + 1], x), x0, x1,
                                  epsabs=0, epsrel=1e-12)[0]
    return tot


def test_sab_xs_bruteforce(oracle):
    """Cross section from the integrator vs independent quadrature of the
    cell-interpolated table to 1e-12: alpha integrals per beta row (log-linear
    interpolant, adaptive quadrature at 1e-12 per cell), trapezoid in beta with
    the first bin scaled by 4/3 (NCSABIntegrator.cc:501-526), xs = sum *
    sigma / (4 E/kT); the last energy puts -E/kT below the beta grid."""
    d, s = _sab_small(oracle)
    alpha, beta, sab = d["alpha"], d["beta"], d["sab"]
    kT = 8.6173303e-5 * d["temperature"]
    for e in (0.003, 0.02, 0.09, 0.6):
        ediv = e / kT
        pts = []
        for ib, b in enumerate(beta):
            if b <= -ediv:
                continue
            kk = ediv + b
            am = max(0.0, kk + ediv - 2 * math.sqrt(ediv * kk))
            ap = kk + ediv + 2 * math.sqrt(ediv * kk)
            pts.append((b, _bruteforce_beta_xs(alpha, sab[ib], am, ap)))
        if -ediv >= beta[0]:
            b0 = -ediv
        else:  # just below the grid (NCSABIntegrator.cc:379-388)
            b0 = min(beta[0] - (beta[1] - beta[0]) * 1e-6, beta[0] - abs(beta[0]) * 1e-13,
                     np.nextafter(beta[0], -np.inf))
        bx = [b0] + [p[0] for p in pts]
        by = [0.0] + [p[1] for p in pts]
        if -ediv >= beta[0]:
            bx[0] = b0 - (bx[1] - b0) / 3.0  # 4/3 first-bin rule
        tot = np.trapezoid(by, bx)
        exact = tot * d["sigma_bound"] / (4 * ediv)
        assert s.analyse_xs(e) == pytest.approx(exact, rel=1e-12), e


def test_sab_xs_provider_limits(oracle):
    d, s = _sab_small(oracle)
    eg, xs = s.egrid()
    # inside: linear interpolation; below: sqrt(E0/E) xs0 (1/v); above: k/E + free gas
    e_mid = 0.5 * (eg[100] + eg[101])
    assert s.xs([e_mid])[0] == pytest.approx(0.5 * (xs[100] + xs[101]), rel=1e-14)
    assert s.xs([eg[0] / 4])[0] == pytest.approx(2.0 * xs[0], rel=1e-14)
    A = d["mass"] / 1.00866491588
    kT = 8.6173303e-5 * d["temperature"]
    sfree = (d["mass"] / (1.00866491588 + d["mass"])) ** 2 * d["sigma_bound"]

    def fg(e):
        a = math.sqrt(A * e / kT)
        return sfree * ((1 + 0.5 / a**2) * math.erf(a) + math.exp(-a * a) / (math.sqrt(math.pi) * a))
    emax = eg[-1]
    k = (xs[-1] - fg(emax)) * emax
    for e in (emax * 1.5, emax * 10):
        assert s.xs([e])[0] == pytest.approx(k / e + fg(e), rel=1e-12)
    # continuity at Emax
    assert s.xs([emax * (1 + 1e-12)])[0] == pytest.approx(xs[-1], rel=1e-9)


def test_sab_sampler_kinematics_and_beta_marginal(oracle):
    """Alg1 samples at the upper energy-grid point and rejects kinematically
    forbidden pairs, with beta piecewise linear between grid rows: its beta
    marginal converges to the exact marginal of the interpolated table as the
    beta grid is refined.  Checked on a 200-row table above the first beta bin."""
    d = synth.small_sab_table(nbeta=200)
    s = oracle.sab(d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    eg, xs = s.egrid()
    kT = 8.6173303e-5 * d["temperature"]
    alpha, beta, sab = d["alpha"], d["beta"], d["sab"]
    for ie, kslim in ((250, 0.01), (290, 0.006)):
        e = float(eg[ie])
        n = 200000
        a, b, nerr = s.sample_ab(np.full(n, e), seed=11)
        assert nerr == 0
        ediv = e / kT
        kk = ediv + b
        assert np.all(kk >= 0)
        am = np.maximum(0.0, kk + ediv - 2 * np.sqrt(ediv * kk))
        ap = kk + ediv + 2 * np.sqrt(ediv * kk)
        assert np.all((a >= am * (1 - 1e-12)) & (a <= ap * (1 + 1e-12)))
        bx, by = [], []
        for ib, bb in enumerate(beta):
            if bb <= -ediv:
                continue
            q = ediv + bb
            lo = max(0.0, q + ediv - 2 * math.sqrt(ediv * q))
            hi = q + ediv + 2 * math.sqrt(ediv * q)
            bx.append(bb)
            by.append(_bruteforce_beta_xs(alpha, sab[ib], lo, hi))
        bx, by = np.array(bx), np.array(by)
        cum = np.concatenate([[0.0], np.cumsum(0.5 * np.diff(bx) * (by[1:] + by[:-1]))])
        cdf = lambda x: np.interp(x, bx, cum / cum[-1])
        ks = stats.kstest(b[b >= bx[0]], cdf).statistic
        assert ks < kslim, (ie, ks)


def test_sab_highe_extender_runs(oracle):
    d, s = _sab_small(oracle)
    eg, _ = s.egrid()
    e = np.geomspace(eg[-1] * 1.01, 5.0, 2000)
    oe, om, nerr = s.sample_iso(e, seed=5)
    assert nerr == 0
    assert np.all(np.abs(om) <= 1.0) and np.all(oe >= 0)


# ---------------------------------------------------------------- SCBragg --
GE_A = 5.65735
GE_D220 = GE_A / math.sqrt(8.0)
GE_220_DEMI = np.array([[2, 2, 0], [2, -2, 0], [2, 0, 2], [2, 0, -2], [0, 2, 2], [0, 2, -2]], dtype=np.int32)


def _ge220(oracle, fsq, mos, prec=1e-3, R=None):
    v0n = GE_A ** 3 * 8
    return oracle.scbragg(GE_A, GE_A, GE_A, v0n, [GE_D220], [fsq], [0, 6], GE_220_DEMI,
                          np.eye(3) if R is None else R, mos, prec)


def test_scbragg_ge220_reference_log(oracle):
    # tests/src/app_backscat/main.cc + test.log: Ge (0 2 -2) family only, mosaicity 1 deg,
    # identity orientation, isotropic average of the single-crystal xs over 1e5 directions
    # ("sgl_cry") and the powder xs ("pc").  |F|^2*mult is fixed by the log's powder value.
    log = [(1.91137, 1.91584), (4.30058, 4.22342), (4.39668, 4.38903), (4.49384, 4.52423),
           (4.59206, 4.53223), (4.69135, 4.66975), (4.79169, 4.91714), (4.8931, 4.97866),
           (4.99557, 5.00076), (5.0991, 5.05824), (5.2037, 5.10551), (5.30925, 4.51393), (0.0, 0.0)]
    thr = 2 * GE_D220
    wls = [thr * 0.6] + list(np.linspace(thr * 0.9, thr * 0.99, 10)) + [thr * 0.99999, thr * 1.00001]
    v0n = GE_A ** 3 * 8
    fm = log[0][0] * 2 * v0n / (wls[0] ** 2 * GE_D220)
    sc = _ge220(oracle, fm / 12, math.pi / 180)
    n = 400000
    dirs = synth.isotropic_dirs(n, seed=11)
    for wl, (pc, sgl) in zip(wls, log):
        if pc:  # the powder value scales as wl^2 for one family
            assert fm * wl * wl * GE_D220 / (2 * v0n) == pytest.approx(pc, rel=5e-6)
        x = sc.xs_dir(np.full(n, WL2EKIN / wl ** 2), dirs, nthreads=8)
        if not sgl:
            assert not x.any()
            continue
        s = x.std()
        pull = (x.mean() - sgl) / math.sqrt(s * s / 1e5 + s * s / n)
        assert abs(pull) < 4.0, (wl, x.mean(), sgl)


def test_scbragg_rocking_curve_closed_form(oracle):
    # one demi-normal in reflection: xs(delta) = Q/(sigma sqrt(2pi)) exp(-delta^2/2sigma^2)
    # with Q = (|F|^2/(V0 n)) * wl^3 / (2 sin(thB) cos(thB)), delta the rocking offset
    v0n = GE_A ** 3 * 8
    fsq = 40.0
    n = np.array([1.0, 1.0, 0.0]) / math.sqrt(2)
    t = np.array([1.0, -1.0, 0.0]) / math.sqrt(2)
    for mos, prec in [(math.pi / 180, 1e-3), (0.1 * math.pi / 180, 1e-4), (20 / 60 * math.pi / 180, 1e-2)]:
        sc = _ge220(oracle, fsq, mos, prec)
        sig = mos * 0.42466090014400952
        for wl in (1.0, 2.0, 3.0):
            tb = math.asin(wl / (2 * GE_D220))
            ds = np.linspace(-3.5, 3.5, 15) * sig
            dirs = np.array([-math.sin(tb + x) * n + math.cos(tb + x) * t for x in ds])
            x = sc.xs_dir(np.full(len(ds), WL2EKIN / wl ** 2), dirs)
            Q = fsq / v0n * wl ** 3 / (2 * math.sin(tb) * math.cos(tb))
            assert x[7] == pytest.approx(Q / (sig * math.sqrt(2 * math.pi)), rel=4e-3)
            np.testing.assert_allclose(x / x[7], np.exp(-ds ** 2 / (2 * sig ** 2)), atol=3e-3)
            # far off the Bragg condition (beyond the truncation) the normal does not reflect
            one = oracle.scbragg(GE_A, GE_A, GE_A, v0n, [GE_D220], [fsq], [0, 1], GE_220_DEMI[:1],
                                 np.eye(3), mos, prec)
            far = -math.sin(tb + 20 * sig) * n + math.cos(tb + 20 * sig) * t
            assert one.xs_dir([WL2EKIN / wl ** 2], far)[0] == 0.0
            assert one.xs_dir([WL2EKIN / wl ** 2], dirs[7])[0] == pytest.approx(x[7], rel=1e-12)


def test_scbragg_isotropic_average_is_powder(oracle):
    # orientation-averaged mosaic crystal = powder (same families, |F|^2 and multiplicities)
    cr = synth.fcc_powder(dcutoff=0.2)
    msd = oracle.debye_msd(410.0, 293.0, 26.98)
    sp = [(cr.species[0].csl, msd, cr.species[0].positions)]
    hkl, mult, dsp, fsq, moff, mh = oracle.hkl_orthogonal_members(cr.a, cr.b, cr.c, sp, cr.dcutoff)
    assert np.array_equal(np.diff(moff) * 2, mult)  # members are demi-normals
    v0n = cr.volume * cr.n_atoms
    th = 0.3
    R = np.array([[1, 0, 0], [0, math.cos(th), -math.sin(th)], [0, math.sin(th), math.cos(th)]])
    sc = oracle.scbragg(cr.a, cr.b, cr.c, v0n, dsp, fsq, moff, mh, R, 30 * math.pi / (180 * 60))
    pb = oracle.powderbragg(v0n, dsp, fsq * mult)
    n = 200000
    dirs = synth.isotropic_dirs(n, seed=3)
    for wl in (1.0, 2.5, 4.0):
        e = WL2EKIN / wl ** 2
        x = sc.xs_dir(np.full(n, e), dirs, nthreads=8)
        assert abs(x.mean() - pb.xs([e])[0]) < 4 * x.std() / math.sqrt(n) + 2e-3 * pb.xs([e])[0]


def test_scbragg_sampling_is_bragg_elastic(oracle):
    cr = synth.fcc_powder(dcutoff=0.3)
    msd = oracle.debye_msd(410.0, 293.0, 26.98)
    sp = [(cr.species[0].csl, msd, cr.species[0].positions)]
    hkl, mult, dsp, fsq, moff, mh = oracle.hkl_orthogonal_members(cr.a, cr.b, cr.c, sp, cr.dcutoff)
    v0n = cr.volume * cr.n_atoms
    sc = oracle.scbragg(cr.a, cr.b, cr.c, v0n, dsp, fsq, moff, mh, np.eye(3), 30 * math.pi / (180 * 60))
    n = 100000
    wl = 2.5
    dirs = synth.isotropic_dirs(n, seed=5)
    e = np.full(n, WL2EKIN / wl ** 2)
    x = sc.xs_dir(e, dirs, nthreads=8)
    oe, od, nerr = sc.sample_dir(e, dirs, seed=7, nthreads=8)
    assert nerr == 0 and np.array_equal(oe, e)
    hit = x > 0
    assert 0.05 < hit.mean() < 0.95
    assert np.array_equal(od[~hit], dirs[~hit])
    assert np.abs(np.linalg.norm(od, axis=1) - 1).max() < 1e-14
    mu = np.sum(od * dirs, axis=1)[hit]
    cand = 1 - wl ** 2 / (2 * dsp[dsp > wl / 2] ** 2)  # cos(2 thB) of the reachable families
    assert np.min(np.abs(mu[:, None] - cand[None, :]), axis=1).max() < 1e-12
