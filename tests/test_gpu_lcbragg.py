"""Layered crystals by NCrystal's reference models on the GPU (the SCBragg
kernels over rotated neutrons, ncamd_create_lcbragg_ref) against the
reference's own pyrolytic-graphite cross sections (lcmode=20 to the printed
6 digits, tests/golden/lc_graphite.json) and the checker
(tests/test_lcbragg_checker.py pins it): cross sections identical to 1e-12,
the random-rotation model converging to the reference's exact model,
sampling elastic with every scattering angle a Bragg angle of a family and
the choice among families distributed as the checker's, neutrons without
cross section unchanged, clones, argument errors; the layered-crystal model
itself (lcmode=0): cross sections against the reference's values and the
checker, scattering against the reference model over 401 rotations."""
import math

import numpy as np
import pytest
from scipy.stats import chi2

from test_lcbragg_checker import FIX, HDR, WL2EKIN, beam, bragg_angles, checker_crystal, graphite_inputs, lchelper

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return core


@pytest.fixture(scope="module")
def crystals(core, oracle):
    def msd(e, d, T, m):
        return core.vdos_analyse((e, d), T, m)["msd"]
    cell, sp, v0n = graphite_inputs(msd)
    hkl, mult, dsp, fsq, moff, mh = core.hkl_spacegroup_members(FIX["spacegroup"], cell, sp, FIX["dcutoff"])
    L = core.reci_lattice(cell)
    g = core.create_scbragg(v0n, dsp, fsq, moff, mh, L, np.eye(3), math.radians(FIX["mosaicity_deg"]))
    o, ofam = checker_crystal(oracle)
    assert np.array_equal(dsp, ofam[2]) and np.max(np.abs(fsq - ofam[3]) / ofam[3]) < 1e-14
    return g, o, dsp


def test_lcmode20_cross_sections(core, oracle, crystals):
    g, o, _ = crystals
    ref = FIX["app_lc"]["20"]
    wl, xs = np.array(ref["wavelength"]), np.array(ref["xs"])
    lc = core.create_lcbragg_ref(g, FIX["lcaxis"], 20)
    _, d = beam(len(wl), 1.0)
    e = WL2EKIN / wl ** 2
    got = lc.crossSection(e, d)
    zero = xs == 0
    assert np.all(got[zero] == 0)
    assert np.max(np.abs(got[~zero] - xs[~zero]) / xs[~zero]) < 6e-6
    want = oracle.LCBraggRef(o, FIX["lcaxis"], 20).xs_dir(e, d)
    assert np.max(np.abs(got - want) / np.maximum(want, 1e-300)) < 1e-12
    # a larger batch with random energies and directions, and a grid of 101
    rng = np.random.default_rng(4)
    n = 3000
    e = WL2EKIN / rng.uniform(0.5, 7.0, n) ** 2
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    for mode in (20, 101):
        got = core.create_lcbragg_ref(g, FIX["lcaxis"], mode).crossSection(e, d)
        want = oracle.LCBraggRef(o, FIX["lcaxis"], mode).xs_dir(e, d)
        den = np.maximum(np.abs(want), 1e-300)
        assert np.max(np.where(want > 0, np.abs(got - want) / den, np.abs(got))) < 1e-12


def test_random_rotations_average_to_the_exact_model(core, crystals):
    g, _, _ = crystals
    ref = FIX["app_lc"]["0"]
    lc = core.create_lcbragg_ref(g, FIX["lcaxis"], -1)
    n = 100000
    for i in (2, 7, 16, 23, 40, 41, 42):
        wl, x0 = ref["wavelength"][i], ref["xs"][i]
        e, d = beam(n, wl)
        x = lc.crossSection(e, d)
        assert abs(x.mean() - x0) < 5 * x.std() / math.sqrt(n) + 2e-3 * x0, (wl, x.mean(), x0)


def _family_counts(od, wl, dsp):
    ang = np.degrees(np.arccos(np.clip(od @ np.array(FIX["direction"]), -1, 1)))
    allowed = bragg_angles(dsp, wl)
    k = np.argmin(np.abs(ang[:, None] - allowed[None, :]), axis=1)
    err = np.abs(ang - allowed[k])
    return np.bincount(k, minlength=len(allowed)), err


def test_sampling(core, oracle, crystals):
    g, o, dsp = crystals
    for mode in (20, -3):
        lc = core.create_lcbragg_ref(g, FIX["lcaxis"], mode)
        olc = oracle.LCBraggRef(o, FIX["lcaxis"], mode)
        for wl in (4.74539, 2.37269, 1.5818):
            n = 40000
            e, d = beam(n, wl)
            lc.setRandomSeed(11)
            ge, gd = lc.sampleScatter(e, d)
            assert np.array_equal(ge, e)
            assert np.allclose(np.linalg.norm(gd, axis=1), 1.0, atol=1e-12)
            gc, gerr = _family_counts(gd, wl, dsp)
            assert np.max(gerr) < 1e-5
            oe, od, nerr = olc.sample_dir(e, d, seed=5)
            assert nerr == 0
            oc, _ = _family_counts(np.asarray(od), wl, dsp)
            m = (gc + oc) > 0
            x2 = float(((gc[m] - oc[m]) ** 2 / (gc[m] + oc[m])).sum())
            p = chi2.sf(x2, max(1, int(m.sum()) - 1))
            print(f"mode {mode} wl {wl}: families {int(m.sum())} p {p:.3g}")
            assert p > 1e-3
    # no cross section at any rotation: unchanged
    lc = core.create_lcbragg_ref(g, FIX["lcaxis"], 20)
    e, d = beam(100, 7.9)
    ge, gd = lc.sampleScatter(e, d)
    assert np.array_equal(ge, e) and np.array_equal(gd, d)


def test_clone_and_device_calls(core, crystals):
    from ncrystal_amd import _capi as C
    g, _, _ = crystals
    lc = core.create_lcbragg_ref(g, FIX["lcaxis"], 20)
    c = core.LCBragg(C.call("ncamd_clone", lc._h, 5), g)
    e, d = beam(500, 3.3)
    assert np.array_equal(c.crossSection(e, d), lc.crossSection(e, d))
    ce, cd = c.sampleScatter(e, d)
    assert np.array_equal(ce, e)


def test_argument_errors(core, crystals):
    from ncrystal_amd import _capi as C
    g, _, _ = crystals
    with pytest.raises(C.NCBadInput):
        core.create_lcbragg_ref(g, FIX["lcaxis"], 0)
    with pytest.raises(C.NCBadInput):
        core.create_lcbragg_ref(g, (0.0, 0.0, 0.0), 20)
    ei = core.create_elincscatter([0.01], [1.0], [1.0])
    with pytest.raises(C.NCBadInput):
        core.create_lcbragg_ref(ei, FIX["lcaxis"], 20)
    lc = core.create_lcbragg_ref(g, FIX["lcaxis"], 20)
    with pytest.raises(C.NCBadInput):
        core.create_composition([lc, ei], [1.0, 1.0])
    with pytest.raises(C.NCBadInput):
        lc.crossSectionIsotropic(np.array([0.01]))
    assert lc.isOriented() if hasattr(lc, "isOriented") else True


def test_lcmode0_cross_sections(core, oracle, crystals):
    """the layered-crystal model (lcmode=0) against the reference's values
    (6 printed digits) and the checker (1e-10: the same plane sets, ranges and
    Romberg steps, device cosines)"""
    def msd(e, d, T, m):
        return core.vdos_analyse((e, d), T, m)["msd"]
    cell, sp, v0n = graphite_inputs(msd)
    hkl, mult, dsp, fsq, moff, mh = core.hkl_spacegroup_members(FIX["spacegroup"], cell, sp, FIX["dcutoff"])
    L = core.reci_lattice(cell)
    lc = core.create_lcbragg(v0n, dsp, fsq, moff, mh, L, np.eye(3), FIX["lcaxis"], math.radians(FIX["mosaicity_deg"]))
    ref = FIX["app_lc"]["0"]
    wl, xs = np.array(ref["wavelength"]), np.array(ref["xs"])
    _, d = beam(len(wl), 1.0)
    got = lc.crossSection(WL2EKIN / wl ** 2, d)
    zero = xs == 0
    assert np.all(got[zero] == 0)
    assert np.max(np.abs(got[~zero] - xs[~zero]) / xs[~zero]) < 6e-6
    o, ofam = checker_crystal(oracle)
    olc = lchelper(oracle, o, ofam)
    rng = np.random.default_rng(8)
    n = 400
    e = WL2EKIN / rng.uniform(0.5, 7.0, n) ** 2
    dd = rng.normal(size=(n, 3))
    dd /= np.linalg.norm(dd, axis=1)[:, None]
    dd[:20] = (0.0, 0.0, 1.0)          # neutrons along the layer axis
    g = lc.crossSection(e, dd)
    w = olc.xs_dir(e, dd)
    assert np.max(np.where(w > 0, np.abs(g - w) / np.where(w > 0, w, 1), np.abs(g))) < 1e-10
    assert np.mean(g > 0) > 0.3
    # the mode != 0 constructor with the same inputs equals the reference-model path
    r20 = core.create_lcbragg(v0n, dsp, fsq, moff, mh, L, np.eye(3), FIX["lcaxis"],
                              math.radians(FIX["mosaicity_deg"]), mode=20)
    assert np.max(np.abs(r20.crossSection(WL2EKIN / wl ** 2, d) - np.array(FIX["app_lc"]["20"]["xs"]))
                  / np.maximum(np.array(FIX["app_lc"]["20"]["xs"]), 1e-300)) < 6e-6


def _graphite_inputs(core):
    def msd(e, d, T, m):
        return core.vdos_analyse((e, d), T, m)["msd"]
    cell, sp, v0n = graphite_inputs(msd)
    hkl, mult, dsp, fsq, moff, mh = core.hkl_spacegroup_members(FIX["spacegroup"], cell, sp, FIX["dcutoff"])
    return v0n, dsp, fsq, moff, mh, core.reci_lattice(cell)


def test_lcmode0_sampling(core, crystals):
    """scattering of the layered-crystal model (lcmode=0) against the
    reference model over 401 rotations (itself compared with the checker
    above): elastic, Bragg angles, the same choice among families and the same
    distribution of the outgoing angle to the layer axis (chi-square), and
    unchanged neutrons without cross section"""
    v0n, dsp, fsq, moff, mh, L = _graphite_inputs(core)
    mos = math.radians(FIX["mosaicity_deg"])
    lc = core.create_lcbragg(v0n, dsp, fsq, moff, mh, L, np.eye(3), FIX["lcaxis"], mos)
    ref = core.create_lcbragg(v0n, dsp, fsq, moff, mh, L, np.eye(3), FIX["lcaxis"], mos, mode=401)
    for wl, direction in ((4.74539, FIX["direction"]), (2.37269, FIX["direction"]), (1.5818, FIX["direction"]),
                          (2.9, (0.3, 0.5, 0.812404))):
        n = 30000
        e = np.full(n, WL2EKIN / wl ** 2)
        d = np.tile(np.asarray(direction) / np.linalg.norm(direction), (n, 1))
        lc.setRandomSeed(21)
        ge, gd = lc.sampleScatter(e, d)
        assert np.array_equal(ge, e)
        assert np.allclose(np.linalg.norm(gd, axis=1), 1.0, atol=1e-12)
        re, rd = ref.sampleScatter(e, d)
        u = d[0]
        for a, b in ((gd, rd),):
            ang_a = np.degrees(np.arccos(np.clip(a @ u, -1, 1)))
            ang_b = np.degrees(np.arccos(np.clip(b @ u, -1, 1)))
            allowed = bragg_angles(dsp, wl)
            ka = np.argmin(np.abs(ang_a[:, None] - allowed[None, :]), axis=1)
            kb = np.argmin(np.abs(ang_b[:, None] - allowed[None, :]), axis=1)
            moved = ang_a > 1e-6
            assert np.max(np.abs(ang_a[moved] - allowed[ka[moved]])) < 1e-5
            ca = np.bincount(ka, minlength=len(allowed))
            cb = np.bincount(kb, minlength=len(allowed))
            m = (ca + cb) > 0
            p1 = chi2.sf(float(((ca[m] - cb[m]) ** 2 / (ca[m] + cb[m])).sum()), max(1, int(m.sum()) - 1))
            ha, _ = np.histogram(a[:, 2], bins=20, range=(-1, 1))
            hb, _ = np.histogram(b[:, 2], bins=20, range=(-1, 1))
            mm = (ha + hb) > 0
            p2 = chi2.sf(float(((ha[mm] - hb[mm]) ** 2 / (ha[mm] + hb[mm])).sum()), max(1, int(mm.sum()) - 1))
            print(f"lcmode 0 vs 401 at {wl} Aa: families p {p1:.3g}, cos to axis p {p2:.3g}")
            assert p1 > 1e-3 and p2 > 1e-3
    e, d = beam(50, 7.9)
    ge, gd = lc.sampleScatter(e, d)
    assert np.array_equal(ge, e) and np.array_equal(gd, d)
