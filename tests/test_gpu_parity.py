"""GPU (HIP, gfx950) results vs the CPU checker on the same seeded inputs,
called through the C ABI.  Tolerances: integer/index results exact,
FP64 cross sections 1e-12 relative, per-neutron samples compared on the same
Philox stream (1e-12 relative except a bounded fraction of branch flips
caused by last-bit libm differences), FP32 PowderBragg cross sections 1e-6
relative to FP64."""
import json
import math

import numpy as np
import pytest
from scipy import stats

from conftest import GOLDEN, assert_streams_match
from ncrystal_amd import synth

pytestmark = pytest.mark.gpu

WL2EKIN = 0.081804209605330899
XS_RTOL = 1e-12


@pytest.fixture(scope="module")
def amd():
    from ncrystal_amd import _capi, core, materials
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return core, materials


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(a), np.abs(b))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(den > 0, np.abs(a - b) / den, 0.0)
    return r


def test_debye_msd(amd, oracle):
    core, _ = amd
    for td, t, m in [(410.0, 293.0, 26.98), (100.0, 0.0, 1.00794), (2000.0, 1000.0, 238.02891)]:
        assert core.debye_msd(td, t, m) == oracle.debye_msd(td, t, m)


@pytest.mark.parametrize("case", ["al_nosg", "fcc_c1", "ortho_small"])
def test_hkl_families(amd, oracle, case):
    core, _ = amd
    if case == "al_nosg":
        g = json.loads((GOLDEN / "nosg_al_hkl.json").read_text())
        m = g["material"]
        msd = oracle.debye_msd(m["debye_temp"], m["temperature"], m["mass"])
        a = b = c = m["a"]
        sp = [(m["csl"], msd, synth._norm_positions(m["positions"]))]
        dcut = m["dcutoff"]
    else:
        cr = synth.fcc_powder() if case == "fcc_c1" else synth.ortho_random(seed=5, dcutoff=0.6)
        a, b, c = cr.a, cr.b, cr.c
        sp = [(s.csl, oracle.debye_msd(s.debye_temp, cr.temperature, s.mass), s.positions) for s in cr.species]
        dcut = cr.dcutoff
    h1, m1, d1, f1 = core.hkl_orthogonal(a, b, c, sp, dcut)
    h2, m2, d2, f2 = oracle.hkl_orthogonal(a, b, c, sp, dcut)
    print(case, "families", len(h1))
    assert len(h1) == len(h2)
    assert np.array_equal(h1, h2)
    assert np.array_equal(m1, m2)
    assert np.array_equal(d1, d2)  # bit-exact d-spacings
    assert rel_err(f1, f2).max() <= 1e-13


def _fcc_pb(core, oracle):
    cr = synth.fcc_powder()
    msd = oracle.debye_msd(410.0, 293.0, 26.98)
    hkl, mult, dsp, fsq = oracle.hkl_orthogonal(cr.a, cr.b, cr.c, [(cr.species[0].csl, msd, cr.species[0].positions)],
                                                cr.dcutoff)
    v0n = cr.volume * cr.n_atoms
    return core.create_powderbragg(v0n, dsp, fsq * mult), oracle.powderbragg(v0n, dsp, fsq * mult)


def test_powderbragg_xs(amd, oracle):
    core, _ = amd
    g, o = _fcc_pb(core, oracle)
    e2d_g, fdm_g, f32 = core.powderbragg_tables(g)
    e2d_o, fdm_o = oracle.pb_tables(o)
    assert np.array_equal(e2d_g, e2d_o) and np.array_equal(fdm_g, fdm_o)
    e = synth.energies_logu(200000, 1e-5, 10.0, seed=1)
    edge = np.concatenate([e2d_o, np.nextafter(e2d_o, 0), np.nextafter(e2d_o, 1), [0.0, -1.0, np.inf, np.nan]])
    e = np.concatenate([e, edge])
    xg = g.crossSectionIsotropic(e)
    xo = o.xs(e)
    assert np.array_equal(np.isnan(xg), np.isnan(xo))
    ok = ~np.isnan(xo)
    assert rel_err(xg[ok], xo[ok]).max() <= XS_RTOL
    # repeat layout of the _many call
    xr = g.crossSectionIsotropic(e[:100], repeat=3)
    assert np.array_equal(xr.reshape(3, 100), np.tile(xg[:100], (3, 1)))
    # survey-recorded reference value at 4 Aa
    assert g.crossSectionIsotropic(WL2EKIN / 16.0) == pytest.approx(1.6308398365212686, rel=1e-15)


def test_powderbragg_fp32(amd, oracle):
    core, _ = amd
    g, o = _fcc_pb(core, oracle)
    e = synth.energies_logu(100000, 1e-5, 10.0, seed=2)
    x64 = g.crossSectionIsotropic(e)
    x32 = g.crossSectionIsotropic_fp32(e)
    nz = x64 > 0
    err = rel_err(x32[nz], x64[nz]).max()
    print("fp32 max rel err", err)
    assert err <= 1e-6
    assert np.array_equal(x32 > 0, nz)


@pytest.mark.parametrize("nplanes", [1, 2, 63, 64, 65, 1023, 1024, 1025, 5000, 262151])
def test_powderbragg_fp32_table(amd, oracle, nplanes):
    """the FP32 cumulative table is the FP64 cumulative rounded to float, to
    within one float ulp, at sizes around the scan's lane and wave counts and
    with contributions spread over ten decades"""
    core, _ = amd
    rng = np.random.default_rng(nplanes)
    d = np.sort(rng.uniform(0.1, 5.0, nplanes))[::-1].copy()
    fm = np.exp(rng.uniform(np.log(1e-6), np.log(1e4), nplanes))
    g = core.create_powderbragg(60.0, d, fm, fp32_tables=True)
    e2d, fdm, f32 = core.powderbragg_tables(g)
    assert f32.dtype == np.float32 and f32.size == fdm.size
    ref = fdm.astype(np.float32)
    ulp = np.spacing(np.abs(ref))
    assert np.all(np.abs(f32.astype(np.float64) - fdm) <= ulp), np.max(np.abs(f32 - fdm) / ulp)


def test_powderbragg_sampling(amd, oracle):
    core, _ = amd
    g, o = _fcc_pb(core, oracle)
    e = synth.energies_logu(50000, 1e-3, 10.0, seed=3)
    g.setRandomSeed(17)
    oe, om = g.sampleScatterIsotropic(e, repeat=2)
    re_, rm, nerr = o.sample_iso(e, repeat=2, seed=17, stream=0)
    assert nerr == 0
    assert np.array_equal(oe, re_)
    assert np.array_equal(om, rm)


def test_powderbragg_c2_ortho(amd, oracle):
    # orthorhombic 24-atom cell, dcutoff 0.1 Aa: ~2.6e5 hkl families, ~7.8e4
    # distinct planes; GPU families/tables vs the oracle, xs bit-exact (FP64)
    # and within 1e-6 (FP32 cumulative tables), sampling identical per neutron
    core, _ = amd
    cr = synth.ortho_random(seed=1, dcutoff=0.1)
    sp = [(s.csl, oracle.debye_msd(s.debye_temp, cr.temperature, s.mass), s.positions) for s in cr.species]
    h1, m1, d1, f1 = core.hkl_orthogonal(cr.a, cr.b, cr.c, sp, cr.dcutoff)
    h2, m2, d2, f2 = oracle.hkl_orthogonal(cr.a, cr.b, cr.c, sp, cr.dcutoff, maxout=3000000)
    print("c2 families", len(h1))
    assert np.array_equal(h1, h2) and np.array_equal(m1, m2) and np.array_equal(d1, d2)
    assert rel_err(f1, f2).max() <= 1e-13
    v0n = cr.volume * cr.n_atoms
    g = core.create_powderbragg(v0n, d2, f2 * m2)
    o = oracle.powderbragg(v0n, d2, f2 * m2)
    ge, gf, g32 = core.powderbragg_tables(g)
    oe_, of_ = oracle.pb_tables(o)
    print("c2 planes", len(ge))
    assert np.array_equal(ge, oe_) and np.array_equal(gf, of_)
    e = synth.energies_logu(1000000, 1e-4, 1.0, seed=3)
    e[:4] = [ge[0], ge[-1], ge[len(ge) // 2], np.nextafter(ge[0], 0)]  # exact thresholds and just below
    x64 = g.crossSectionIsotropic(e)
    assert np.array_equal(x64, o.xs(e, nthreads=8))
    x32 = g.crossSectionIsotropic_fp32(e)
    nz = x64 > 0
    err = rel_err(x32[nz], x64[nz]).max()
    print("c2 fp32 max rel err", err)
    assert err <= 1e-6 and np.array_equal(x32 > 0, nz)
    g.setRandomSeed(5)
    se, sm = g.sampleScatterIsotropic(e)
    re_, rm, nerr = o.sample_iso(e, seed=5, nthreads=8)
    assert nerr == 0 and np.array_equal(se, re_) and np.array_equal(sm, rm)


def test_elinc(amd, oracle):
    core, _ = amd
    msd = [0.01, 0.02, 0.005]
    bixs = [2.0, 0.3, 1.1]
    scale = [0.5, 0.25, 0.25]
    g = core.create_elincscatter(msd, bixs, scale)
    o = oracle.elinc(msd, bixs, scale)
    e = synth.energies_logu(100000, 1e-6, 10.0, seed=4)
    assert rel_err(g.crossSectionIsotropic(e), o.xs(e)).max() <= XS_RTOL
    g.setRandomSeed(5)
    oe, om = g.sampleScatterIsotropic(e)
    re_, rm, _ = o.sample_iso(e, seed=5)
    assert np.array_equal(oe, re_)
    r = rel_err(om, rm)
    assert np.mean(r > 1e-12) < 1e-3, np.mean(r > 1e-12)


def _small_sab(core, oracle, **kw):
    d = synth.small_sab_table(**kw)
    args = (d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    return d, core.create_sabscatter(*args), oracle.sab(*args)


def test_sab_tables(amd, oracle):
    core, _ = amd
    d, g, o = _small_sab(core, oracle)
    na, nb = len(d["alpha"]), len(d["beta"])
    lg, cg = g.derived(na, nb)
    lo, co = o.derived(na, nb)
    fin = np.isfinite(lo)
    assert np.array_equal(np.isfinite(lg), fin)
    assert rel_err(lg[fin], lo[fin]).max() <= 1e-14
    assert rel_err(cg, co).max() <= 1e-13
    eg_g, xs_g = g.egrid()
    eg_o, xs_o = o.egrid()
    assert np.array_equal(eg_g, eg_o)
    assert rel_err(xs_g, xs_o).max() <= XS_RTOL
    for ie in (0, 17, 150, len(eg_o) - 1):
        sg = g.sampler(ie, nb)
        so = o.sampler(ie)
        for k in ("kind", "ibeta_low", "nvals"):
            assert sg[k] == so[k], (ie, k)
        assert sg["firstbin"] == so["firstbin"]
        if so["kind"] == 1:
            for k in ("bx", "by", "bcdf", "info"):
                assert rel_err(sg[k], so[k]).max() <= 1e-12, (ie, k)
            assert np.array_equal(sg["infoidx"], so["infoidx"])


def test_sab_xs(amd, oracle):
    core, _ = amd
    d, g, o = _small_sab(core, oracle)
    e = np.concatenate([synth.energies_logu(100000, 1e-7, 20.0, seed=6), g.egrid()[0]])
    assert rel_err(g.crossSectionIsotropic(e), o.xs(e)).max() <= XS_RTOL


def _compare_samples(oe, om, re_, rm, label):
    # same Philox streams: identical except where a last-ulp libm difference
    # flips a branch; such neutrons are rare and counted
    r = np.maximum(rel_err(oe, re_), rel_err(om, rm))
    print(label, "median rel", np.median(r))
    return assert_streams_match(r > 1e-9, label)


def test_sab_sampling(amd, oracle):
    core, _ = amd
    d, g, o = _small_sab(core, oracle)
    e = synth.energies_logu(100000, 1e-5, 5.0, seed=8)  # includes the free-gas extender above Emax
    g.setRandomSeed(99)
    oe, om = g.sampleScatterIsotropic(e)
    re_, rm, nerr = o.sample_iso(e, seed=99)
    assert nerr == 0
    assert np.all(np.abs(om) <= 1.0) and np.all(oe >= 0.0)
    _compare_samples(oe, om, re_, rm, "sab")
    for x, y in ((oe, re_), (om, rm)):
        assert stats.ks_2samp(x, y).statistic < 0.01
    # repeat layout through the split main/high-energy kernels
    g.setRandomSeed(7)
    oe2, om2 = g.sampleScatterIsotropic(e[:20000], repeat=3)
    re2, rm2, nerr = o.sample_iso(e[:20000], repeat=3, seed=7)
    assert nerr == 0
    _compare_samples(oe2, om2, re2, rm2, "sab repeat")


def test_sab_c3_table(amd, oracle):
    core, _ = amd
    d = synth.sab_table_c3()
    args = (d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    g = core.create_sabscatter(*args)
    o = oracle.sab(*args)
    eg_g, xs_g = g.egrid()
    eg_o, xs_o = o.egrid()
    assert np.array_equal(eg_g, eg_o)
    assert rel_err(xs_g, xs_o).max() <= XS_RTOL
    e = synth.energies_logu(50000, 1e-3, 1.0, seed=9)
    assert rel_err(g.crossSectionIsotropic(e), o.xs(e)).max() <= XS_RTOL
    g.setRandomSeed(3)
    oe, om = g.sampleScatterIsotropic(e)
    re_, rm, nerr = o.sample_iso(e, seed=3, nthreads=8)
    assert nerr == 0
    _compare_samples(oe, om, re_, rm, "c3")


def test_composition_c1(amd, oracle):
    """the fcc powder workload: ElInc + PowderBragg + S(alpha,beta) of the
    Debye VDOS (rho ~ E^2 to k_B 410 K, 293 K), 1e6 neutrons with E
    log-uniform 1e-5..10 eV (about a fifth above the table's Emax, in the
    free-gas extender)"""
    core, materials = amd
    cr = synth.fcc_powder()
    sabd = synth.sab_table_c1()
    m = materials.Material(cr, sab=sabd)
    # oracle composition with the same inputs and order (ElInc, Bragg, SAB)
    msd = m.msd
    mm, bb, ss = materials.elinc_inputs(cr, msd)
    v0n, dsp, fm = m.pb_inputs
    o_ei = oracle.elinc(mm, bb, ss)
    o_pb = oracle.powderbragg(v0n, dsp, fm)
    o_sab = oracle.sab(sabd["alpha"], sabd["beta"], sabd["sab"], sabd["temperature"], sabd["mass"],
                       sabd["sigma_bound"])
    comp = oracle.composition([o_ei, o_pb, o_sab], [1.0, 1.0, 1.0])
    e = synth.energies_logu(1000000, 1e-5, 10.0, seed=1)
    emax = o_sab.egrid()[0][-1]
    print("C1: fraction above Emax", np.mean(e >= emax))
    assert rel_err(m.scatter.crossSectionIsotropic(e), comp.xs(e, nthreads=8)).max() <= XS_RTOL
    m.scatter.setRandomSeed(1)
    oe, om = m.scatter.sampleScatterIsotropic(e)
    re_, rm, nerr = comp.sample_iso(e, seed=1, nthreads=8)
    assert nerr == 0
    _compare_samples(oe, om, re_, rm, "composition")
    # oriented interface on an isotropic material: |dir|=1 and mu = dir_in.dir_out
    dirs = synth.isotropic_dirs(len(e), seed=4)
    m.scatter.setRandomSeed(2)
    oe2, od = m.scatter.sampleScatter(e, dirs)
    assert np.allclose(np.linalg.norm(od, axis=1), 1.0, atol=1e-12)
    re2, rd, nerr = comp.sample_dir(e, dirs, seed=2, nthreads=8)
    assert nerr == 0
    assert_streams_match(np.max(np.abs(od - rd), axis=1) > 1e-9, "oriented composition dirs")


def _chi2_2sample(x1, y1, x2, y2, nb=40):
    """two-sample chi-square on a 2-D histogram with equal-population bins of
    the pooled sample in x, then in y within each x slice"""
    xe = np.quantile(np.concatenate([x1, x2]), np.linspace(0, 1, nb + 1))
    xe[0], xe[-1] = -np.inf, np.inf
    chi2, dof = 0.0, -1
    for i in range(nb):
        s1 = (x1 >= xe[i]) & (x1 < xe[i + 1])
        s2 = (x2 >= xe[i]) & (x2 < xe[i + 1])
        ye = np.quantile(np.concatenate([y1[s1], y2[s2]]), np.linspace(0, 1, nb + 1))
        ye[0], ye[-1] = -np.inf, np.inf
        h1 = np.histogram(y1[s1], ye)[0].astype(float)
        h2 = np.histogram(y2[s2], ye)[0].astype(float)
        k1, k2 = math.sqrt(len(x2) / len(x1)), math.sqrt(len(x1) / len(x2))
        m = (h1 + h2) > 0
        chi2 += np.sum((k1 * h1[m] - k2 * h2[m]) ** 2 / (h1[m] + h2[m]))
        dof += int(m.sum())
    return chi2, dof


@pytest.mark.parametrize("E", [0.001, 0.025, 0.5])
def test_sab_alphabeta_chi2_c3(amd, oracle, E):
    """(alpha, beta) of the plain GPU sampler on the 200 x 800 table against
    the CPU checker's sampler on independent streams: two-sample chi-square
    of the 2-D histogram (1600 cells) at 1e7 samples each, and KS of each
    marginal"""
    from ncrystal_amd import _capi as C
    core, _ = amd
    d = synth.sab_table_c3(seed=1)
    g = core.create_sabscatter(d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    o = oracle.sab(d["alpha"], d["beta"], d["sab"], d["temperature"], d["mass"], d["sigma_bound"])
    n = 10_000_000
    st = C.Stream()
    de = C.DeviceArray.from_numpy(np.full(n, E))
    da, db = C.DeviceArray(n), C.DeviceArray(n)
    g.setRandomSeed(4242)
    g.dev_sample_alphabeta(de.ptr, n, da.ptr, db.ptr, stream=st.ptr)
    assert g.dev_check_errors(stream=st.ptr) == 0
    ga, gb = da.numpy(), db.numpy()
    oa, ob, nerr = o.sample_ab(np.full(n, E), seed=99, nthreads=8)
    assert nerr == 0
    chi2, dof = _chi2_2sample(gb, ga, ob, oa)
    p = stats.chi2.sf(chi2, dof)
    print(f"E={E}: chi2 {chi2:.1f} for {dof} dof, p = {p:.3g}")
    assert p > 1e-4
    lim = 1.95 * math.sqrt(2.0 / n)
    assert stats.ks_2samp(ga, oa).statistic < lim and stats.ks_2samp(gb, ob).statistic < lim


def test_rng_state_roundtrip(amd, oracle):
    core, _ = amd
    g, o = _fcc_pb(core, oracle)
    e = synth.energies_logu(1000, 1e-2, 1.0, seed=1)
    g.setRandomSeed(123)
    a = g.sampleScatterIsotropic(e)[1]
    st = g.getRNGState()
    b = g.sampleScatterIsotropic(e)[1]
    g.setRNGState(*st)
    c = g.sampleScatterIsotropic(e)[1]
    assert np.array_equal(b, c) and not np.array_equal(a, b)


def test_dev_pointer_api(amd, oracle):
    from ncrystal_amd import _capi as C
    core, _ = amd
    d, g, o = _small_sab(core, oracle)
    e = synth.energies_logu(100000, 1e-4, 1.0, seed=12)
    st = C.Stream()
    de = C.DeviceArray.from_numpy(e)
    xs, oe, om = C.DeviceArray(len(e)), C.DeviceArray(len(e)), C.DeviceArray(len(e))
    g.dev_crossSectionIsotropic(de.ptr, len(e), xs.ptr, stream=st.ptr)
    g.setRandomSeed(8)
    g.dev_sampleScatterIsotropic(de.ptr, len(e), oe.ptr, om.ptr, stream=st.ptr)
    assert g.dev_check_errors(stream=st.ptr) == 0
    st.sync()
    assert rel_err(xs.numpy(), o.xs(e)).max() <= XS_RTOL
    re_, rm, _ = o.sample_iso(e, seed=8)
    _compare_samples(oe.numpy(), om.numpy(), re_, rm, "dev")
    # device-side energy generator: log-uniform in range
    C.call("ncamd_dev_fill_loguniform", de.ptr, len(e), 1e-3, 1.0, 5, st.ptr)
    st.sync()
    v = de.numpy()
    assert v.min() >= 1e-3 and v.max() <= 1.0
    assert stats.kstest(np.log(v), "uniform", args=(math.log(1e-3), -math.log(1e-3))).pvalue > 1e-4
