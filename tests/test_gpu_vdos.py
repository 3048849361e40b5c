"""VDOS -> S(alpha,beta) on the GPU (ncamd_vdos_kernel: host chain of
regularisation, VDOSEval, G_n convolutions, orders and grids; GPU operand
tables and the FP64 matrix-core sum over orders) against the CPU checker,
whose kernels tests/test_vdos_pins.py pins to the reference's logs:

* grids, number of orders, suggested Emax and MSD identical; S(alpha,beta)
  within 1e-12 relative wherever either value exceeds 1e-250 (the GPU sums
  the orders in another order and evaluates exp on the device);
* cases: Debye 400 K at 300 K (vdoslux 1), the same at 0.5 K (zero edges
  trimmed in the standard form), target Emax 50 eV (1123 orders, of which
  the reference's block split sums 1120), aluminium and polyethylene H of
  the data library at 293.15 K (vdoslux 3 and 2);
* the SABScatter made from it (ncamd_create_sabscatter_from_vdos) against
  the checker's SAB of the same kernel: cross sections within 1e-10,
  samples on the same Philox streams;
* tests/scripts/basic.log through the library: aluminium PowderBragg at
  1-4 Aa with the MSD of its VDOS, absorption, and incoherent elastic plus
  the VDOS S(alpha,beta) at 4 Aa;
* the fcc powder workload (ElInc + PowderBragg + Debye-VDOS S(alpha,beta))
  against the checker's composition with its own expansion;
* CalcError (with the reference's message) for an Emax the expansion
  cannot reach, BadInput for a bad vdoslux."""
import json
import math
import pathlib

import numpy as np
import pytest

from ncrystal_amd import synth
from conftest import assert_streams_match

pytestmark = pytest.mark.gpu

FIX = json.loads((pathlib.Path(__file__).parent / "golden" / "vdos_reference.json").read_text())


@pytest.fixture(scope="module")
def amd():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return _capi, core


def _case(name):
    if name.startswith("debye"):
        return 400.0
    v = FIX["vdos"][name]
    return np.array(v["egrid"]), np.array(v["density"])


CASES = [
    ("debye", 300.0, 27.0, 1, 0.0),
    ("debye", 0.5, 27.0, 1, 0.0),
    ("debye", 300.0, 27.0, 1, 50.0),
    ("Al_sg225/Al", 293.15, 26.981538408, 3, 0.0),
    ("Polyethylene_CH2/H", 293.15, 1.00794, 2, 0.0),
]


def _checker_kernel(oracle, vin, T, m, lux, temax):
    if np.isscalar(vin):
        lo, hi, d = oracle.vdos_debye(vin)
        lo, hi, d = oracle.vdos_regularise(np.array([lo, hi]), d)
    else:
        lo, hi, d = oracle.vdos_regularise(*vin)
    return oracle.vdos_kernel(lo, hi, d, T, m, vdoslux=lux, target_emax=temax)


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-T{c[1]}-lux{c[3]}-E{c[4]}" for c in CASES])
def test_kernel_matches_checker(amd, oracle, case):
    C, core = amd
    name, T, m, lux, temax = case
    if name not in ("debye",) and name not in FIX["vdos"]:
        pytest.skip(f"{name} not in the fixture set")
    vin = _case(name)
    g = core.vdos_kernel(vin, T, m, vdoslux=lux, target_emax=temax)
    o = _checker_kernel(oracle, vin, T, m, lux, temax)
    assert np.array_equal(g["alpha"], o["alpha"]) and np.array_equal(g["beta"], o["beta"])
    assert g["norders"] == o["norders"] and g["suggested_emax"] == o["suggested_emax"] and g["msd"] == o["msd"]
    gs, os_ = g["sab"], o["sab"]
    big = np.maximum(np.abs(gs), np.abs(os_)) > 1e-250
    rel = np.abs(gs - os_)[big] / np.maximum(np.abs(gs), np.abs(os_))[big]
    print(f"{name} T={T} lux={lux}: {g['alpha'].size}x{g['beta'].size}, {g['norders']} orders, "
          f"max rel diff {rel.max():.2e}")
    if rel.max() > 1e-12:
        full = np.zeros_like(gs)
        full[big] = rel
        for flat in np.argsort(full.ravel())[::-1][:8]:
            j, i = divmod(int(flat), gs.shape[1])
            print(f"  row {j} (beta {g['beta'][j]:.6g}) col {i} (alpha {g['alpha'][i]:.6g}): "
                  f"gpu {gs[j, i]:.17g} checker {os_[j, i]:.17g}")
        print("  cells beyond 1e-12:", int((full > 1e-12).sum()), "of", int(big.sum()))
    assert rel.max() <= 1e-12
    assert np.all(np.abs(gs[~big]) <= 1e-240)
    assert np.all(gs >= 0.0)


@pytest.mark.parametrize("name,T,m", [("Al_sg225/Al", 293.15, 26.981538408), ("Polyethylene_CH2/H", 293.15, 1.00794),
                                      ("debye", 300.0, 27.0)])
def test_gn_bit_identical(amd, oracle, name, T, m):
    """G_n of the library (large transforms on the GPU, generations of
    orders on host threads) equal to the checker's one-thread host chain bit
    for bit: the truncation and thinning edges of every higher order, and so
    the number of orders and the grids, follow from these values"""
    C, core = amd
    vin = _case(name)
    if np.isscalar(vin):
        lo, hi, d = oracle.vdos_debye(vin)
        lo, hi, d = oracle.vdos_regularise(np.array([lo, hi]), d)
    else:
        lo, hi, d = oracle.vdos_regularise(*vin)
    for order in (1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 17, 33, 40, 100):
        eg, sg = core.vdos_gn(vin, T, m, order)
        eo, so = oracle.vdos_gn(lo, hi, d, T, m, order)
        print(f"{name} order {order}: {sg.size} points")
        assert np.array_equal(eg, eo) and np.array_equal(sg, so), f"order {order}"


def test_sabscatter_from_vdos_matches_checker(amd, oracle):
    C, core = amd
    vin = _case("Al_sg225/Al")
    T, m = 293.15, 26.981538408
    sigma = 4 * math.pi * 0.3449 ** 2 + 0.0082
    g = core.create_sabscatter_from_vdos(vin, T, m, sigma, vdoslux=2)
    k = _checker_kernel(oracle, vin, T, m, 2, 0.0)
    sf = oracle.sab_std_format(0, k["alpha"], k["beta"], k["sab"], beta_grid_optimised=True)
    o = oracle.sab(sf["alpha"], sf["beta"], sf["sab"], T, m, sigma, suggested_emax=k["suggested_emax"])
    e = synth.energies_logu(200000, 1e-5, 20.0, seed=4)
    xg, xo = g.crossSectionIsotropic(e), o.xs(e, nthreads=8)
    rel = np.abs(xg - xo) / np.maximum(np.abs(xo), 1e-300)
    print("xs max rel diff", rel.max())
    assert rel.max() <= 1e-10
    g.setRandomSeed(12)
    ge, gm = g.sampleScatterIsotropic(e)
    oe, om, nerr = o.sample_iso(e, seed=12, nthreads=8)
    assert nerr == 0
    assert_streams_match((np.abs(ge - oe) > 1e-9 * np.abs(oe)) | (np.abs(gm - om) > 1e-9), "vdos sab iso")


def test_basic_log_aluminium(amd):
    C, core = amd
    b = FIX["basic"]
    v = FIX["vdos"][b["vdos"]]
    vin = (np.array(v["egrid"]), np.array(v["density"]))
    T, mass = b["temperature"], b["mass"]
    msd = core.vdos_analyse(vin, T, mass)["msd"]
    a = b["a"]
    hkl, mult, dsp, fsq, _, _ = core.hkl_spacegroup_members(b["spacegroup"], (a, a, a, 90, 90, 90),
                                                             [(b["coh_sl"], msd, np.array(b["positions"]))],
                                                             b["dcutoff"])
    wl = np.array([1.0, 2.0, 3.0, 4.0])
    e = 0.081804209605330899 / wl ** 2
    pb = core.create_powderbragg(a ** 3 * len(b["positions"]), dsp, fsq * mult).crossSectionIsotropic(e)
    assert np.allclose(pb, b["powderbragg_1to4"], rtol=5e-6, atol=0), pb
    ab = core.create_absoov(b["sigma_abs"]).crossSectionIsotropic(e[3:])[0]
    assert abs(ab / b["absorption_4"] - 1) < 5e-6
    sigma = 4 * math.pi * b["coh_sl"] ** 2 + b["sigma_inc"]
    inel = core.create_sabscatter_from_vdos(vin, T, mass, sigma, vdoslux=3)
    ei = core.create_elincscatter([msd], [b["sigma_inc"]], [1.0])
    bkgd = core.create_composition([ei, inel], [1.0, 1.0]).crossSectionIsotropic(e[3:])[0]
    print("Al 4 Aa: Bragg", pb[3], "bkgd", bkgd)
    assert abs(bkgd / b["bkgd_4"] - 1) < 5e-6
    assert abs((bkgd + pb[3]) / b["bkgd_plus_bragg_4"] - 1) < 5e-6


def test_errors(amd):
    C, core = amd
    err = [r for r in FIX["kernel"] if "error" in r][0]
    with pytest.raises(C.NCCalcError) as exc:
        core.vdos_kernel(err["debye_temp"], err["temperature"], err["mass"], vdoslux=err["vdoslux"],
                         target_emax=err["target_emax"])
    assert str(exc.value) == err["error"] or err["error"] in str(exc.value)
    with pytest.raises(C.NCBadInput):
        core.vdos_kernel(400.0, 300.0, 27.0, vdoslux=6)


def test_c1_material_matches_checker(amd, oracle):
    """the fcc powder workload as NCrystal builds it from a Debye temperature:
    ElInc + PowderBragg + the S(alpha,beta) expanded from the Debye VDOS
    (vdoslux 3), against the checker's composition with its own expansion:
    cross sections within 1e-10, oriented samples on the same streams"""
    C, core = amd
    from ncrystal_amd import materials
    cr = synth.fcc_powder()
    sp = cr.species[0]
    sigma = 4 * math.pi * sp.csl ** 2 + sp.sigma_inc
    m = materials.Material(cr, dynamics=[(1.0, {"vdos": sp.debye_temp, "mass": sp.mass, "sigma_bound": sigma,
                                                "vdoslux": 3})])
    k = _checker_kernel(oracle, sp.debye_temp, cr.temperature, sp.mass, 3, 0.0)
    sf = oracle.sab_std_format(0, k["alpha"], k["beta"], k["sab"], beta_grid_optimised=True)
    mm, bb, ss = materials.elinc_inputs(cr, m.msd)
    o = oracle.composition([oracle.elinc(mm, bb, ss), oracle.powderbragg(*m.pb_inputs),
                            oracle.sab(sf["alpha"], sf["beta"], sf["sab"], cr.temperature, sp.mass, sigma,
                                       suggested_emax=k["suggested_emax"])], [1.0, 1.0, 1.0])
    n = 200000
    e = synth.energies_logu(n, 1e-5, 10.0, seed=1)
    xg, xo = m.scatter.crossSectionIsotropic(e), o.xs(e, nthreads=8)
    assert np.max(np.abs(xg - xo) / np.maximum(xo, 1e-300)) <= 1e-10
    dirs = synth.isotropic_dirs(n, seed=1)
    m.scatter.setRandomSeed(5)
    ge, gd = m.scatter.sampleScatter(e, dirs)
    oe, od, nerr = o.sample_dir(e, dirs, seed=5, nthreads=8)
    assert nerr == 0
    assert_streams_match((np.abs(ge - oe) > 1e-9 * np.abs(oe)) | (np.max(np.abs(gd - od), axis=1) > 1e-9),
                         "C1 with the Debye VDOS expansion")
