"""Layered crystals by NCrystal's reference models (LCBraggRef and
LCBraggRndmRot, ncrystal_core/src/extd_utils/NCLCRefModels.cc) in the
checker (oracle.pyoracle.LCBraggRef over the checker's SCBragg), pinned by
the reference's own outputs for pyrolytic graphite (tests/golden/
lc_graphite.json, from tests/src/app_lc/test.log and
tests/scripts/ncmat2hkl.log):

* the material: hkl families (hkl, multiplicity, d and |F|^2 to the printed
  digits) from space group 194 with the data library's VDOS for the MSD;
* lcmode=20 (LCBraggRef, 23 equally spaced rotations): the cross sections at
  43 wavelengths to the printed 6 digits;
* lcmode=0 (the reference's exact model, the continuous average over
  rotations): a fine grid of rotations reproduces its cross sections, and
  the mean of many single random rotations (lcmode=-1) converges to them;
* sampling: elastic, and every scattering angle is 2 asin(lambda / 2d) of a
  family that can reflect (at 4.74539 Aa only (002): 90 degrees);
* the layered-crystal model itself (lcmode=0, oracle.pyoracle.LCHelperXS:
  plane sets, rotation ranges, Romberg averages): its cross sections at the
  43 wavelengths to the printed 6 digits.
"""
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN

FIX = json.loads((GOLDEN / "lc_graphite.json").read_text())
VDOS = json.loads((GOLDEN / "vdos_reference.json").read_text())["vdos"]["C_sg194_pyrolytic_graphite/C"]
HDR = FIX["laz"]["header"]
WL2EKIN = 0.081804209605330899


def graphite_inputs(vdos_analyse_msd):
    """(cell, species, v0n) of the test's graphite; the MSD from a function
    of (egrid, density, T, mass)"""
    a, c = HDR["lattice_a"], HDR["lattice_c"]
    msd = vdos_analyse_msd(np.array(VDOS["egrid"]), np.array(VDOS["density"]), HDR["temperature"],
                           HDR["average_mass"])
    csl = math.sqrt(HDR["sigma_coh"] / (4 * math.pi))
    return (a, a, c, 90, 90, 120), [(csl, msd, np.array(FIX["positions"]))], HDR["Vc"] * len(FIX["positions"])


def checker_crystal(oracle, dcutoff=None):
    def msd(e, d, T, m):
        lo, hi, dd = oracle.vdos_regularise(e, d)
        return oracle.vdos_analyse(lo, hi, dd, T, m)["msd"]
    cell, sp, v0n = graphite_inputs(msd)
    fam = oracle.hkl_spacegroup_members(FIX["spacegroup"], cell, sp, dcutoff or FIX["dcutoff"])
    hkl, mult, dsp, fsq, moff, mh = fam
    L = oracle.reci_lattice(cell)
    # dir1 = (001) normal -> lab z, dir2 = (010) normal -> lab y: both normals
    # already lie along z and y in the crystal frame, so crystal = lab
    sc = oracle.scbragg(cell[0], cell[1], cell[2], v0n, dsp, fsq, moff, mh, np.eye(3),
                        math.radians(FIX["mosaicity_deg"]), reci=L)
    return sc, fam


def beam(n, wl):
    e = np.full(n, WL2EKIN / wl ** 2)
    d = np.tile(np.array(FIX["direction"]), (n, 1))
    return e, d


def test_graphite_families(oracle):
    _, (hkl, mult, dsp, fsq, _, _) = checker_crystal(oracle)
    rows = np.array(FIX["laz"]["table"])
    n = len(rows)
    assert np.array_equal(hkl[:n], rows[:, :3].astype(int))
    assert np.array_equal(mult[:n], rows[:, 3].astype(int))
    assert np.max(np.abs(dsp[:n] - rows[:, 4]) / rows[:, 4]) < 1e-9
    assert np.max(np.abs(fsq[:n] - rows[:, 5]) / rows[:, 5]) < 1e-9


@pytest.fixture(scope="module")
def crystal(oracle):
    return checker_crystal(oracle)[0]


def test_lcmode20_cross_sections(oracle, crystal):
    ref = FIX["app_lc"]["20"]
    wl, xs = np.array(ref["wavelength"]), np.array(ref["xs"])
    lc = oracle.LCBraggRef(crystal, FIX["lcaxis"], 20)
    assert lc.nprime == 23
    e, d = beam(len(wl), 1.0)
    got = lc.xs_dir(WL2EKIN / wl ** 2, d)
    zero = xs == 0
    assert np.all(got[zero] == 0)
    assert np.max(np.abs(got[~zero] - xs[~zero]) / xs[~zero]) < 6e-6   # 6 printed digits


def test_fine_grid_is_the_exact_model(oracle, crystal):
    """the reference's lcmode=0 values are the exact rotational average (to
    its integration precision, 1e-3); a 401-point grid approaches them"""
    ref = FIX["app_lc"]["0"]
    wl, xs = np.array(ref["wavelength"]), np.array(ref["xs"])
    lc = oracle.LCBraggRef(crystal, FIX["lcaxis"], 401)
    _, d = beam(len(wl), 1.0)
    got = lc.xs_dir(WL2EKIN / wl ** 2, d)
    big = xs > 0.01
    assert np.max(np.abs(got[big] - xs[big]) / xs[big]) < 2e-3
    assert np.all(got[xs == 0] == 0)


def test_random_rotations_average_to_the_exact_model(oracle, crystal):
    ref = FIX["app_lc"]["0"]
    lc = oracle.LCBraggRef(crystal, FIX["lcaxis"], -1)
    n = 20000
    for i in (2, 7, 16, 23, 40, 41, 42):
        wl, x0 = ref["wavelength"][i], ref["xs"][i]
        e, d = beam(n, wl)
        x = lc.xs_dir(e, d, seed=i)
        assert abs(x.mean() - x0) < 5 * x.std() / math.sqrt(n) + 2e-3 * x0, (wl, x.mean(), x0)


def bragg_angles(dsp, wl):
    ok = dsp >= wl / 2
    return np.degrees(2 * np.arcsin(wl / (2 * dsp[ok])))


def test_sampling_angles(oracle, crystal):
    _, (_, _, dsp, _, _, _) = checker_crystal(oracle)
    lc = oracle.LCBraggRef(crystal, FIX["lcaxis"], 20)
    for wl in (4.74539, 2.37269, 1.5818):
        e, d = beam(4000, wl)
        oe, od, nerr = lc.sample_dir(e, d, seed=3)
        assert nerr == 0 and np.array_equal(oe, e)
        assert np.allclose(np.linalg.norm(od, axis=1), 1.0, atol=1e-12)
        ang = np.degrees(np.arccos(np.clip(od @ np.array(FIX["direction"]), -1, 1)))
        moved = ang > 1e-6
        assert moved.mean() > 0.99
        allowed = bragg_angles(dsp, wl)
        err = np.min(np.abs(ang[moved][:, None] - allowed[None, :]), axis=1)
        assert np.max(err) < 1e-5, wl
        if wl == 4.74539:
            assert np.max(np.abs(ang - 90.0)) < 1e-3


def lchelper(oracle, crystal, fam):
    hkl, mult, dsp, fsq, moff, mh = fam
    L = oracle.reci_lattice((HDR["lattice_a"], HDR["lattice_a"], HDR["lattice_c"], 90, 90, 120)).reshape(3, 3)
    nrm = mh @ L.T
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    of = np.repeat(np.arange(len(dsp)), np.diff(moff))
    return oracle.LCHelperXS(crystal, nrm, fsq[of], dsp[of], FIX["lcaxis"], FIX["lcaxis"],
                             HDR["Vc"] * len(FIX["positions"]))


def test_lcmode0_cross_sections(oracle):
    crystal, fam = checker_crystal(oracle)
    lc = lchelper(oracle, crystal, fam)
    ref = FIX["app_lc"]["0"]
    wl, xs = np.array(ref["wavelength"]), np.array(ref["xs"])
    _, d = beam(len(wl), 1.0)
    got = lc.xs_dir(WL2EKIN / wl ** 2, d)
    zero = xs == 0
    assert np.all(got[zero] == 0)
    assert np.max(np.abs(got[~zero] - xs[~zero]) / xs[~zero]) < 6e-6
