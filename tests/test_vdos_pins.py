"""Pins of the checker's VDOS -> S(alpha,beta) expansion (oracle/ncvdos.c)
against what the reference's own tests record for its data library
(tests/golden/vdos_reference.json, written by make_vdos_fixtures.py):

* regulariseVDOSGrid: grid ends, point count and the first and last three
  densities of tests/scripts/vdosregularise.log (73 densities);
* VDOSEval: gamma0, the integral of the given density, the MSD and the
  effective temperature of tests/scripts/analysevdos.log (35 densities at
  1e-3, 300 and 1e6 K; the log prints 12 digits);
* VDOSGn: integral, mean, rms and standardised moments 3-7 of G1, G2, G5,
  G8, G20 and G50 of tests/scripts/vdosgn.log (Au, polyethylene H and C,
  LiH, graphite at 10 and 600 K; 7 digits), computed as the reference's
  moment utility does (the piece intercept in x against powers of x - mean);
* createScatteringKernel: grid sizes, suggested Emax, trimmed edges and the
  "too slow" error of tests/scripts/vdos2knl.log;
* the MSD of a Debye VDOS against the Debye model's closed form
  (ncrystal_core/src/phys_utils/NCDebyeMSD.cc, itself pinned to
  app_debyemsd);
* tests/scripts/basic.log: the aluminium cross sections that depend on
  the VDOS (PowderBragg through its MSD, incoherent elastic + inelastic
  through the expanded kernel), and absorption;
* the Sjolander fill: S at sampled cells against an independent sum over
  orders with x^n exp(-x)/n! from lgamma and the checker's G_n, and detailed
  balance S(alpha, +beta) = S(alpha, -beta) exp(-beta) on the mirrored rows.
CPU only."""
import json
import math
import pathlib

import numpy as np
import pytest

FIX = json.loads((pathlib.Path(__file__).parent / "golden" / "vdos_reference.json").read_text())
KB = 8.6173303e-5


def _vdos(oracle, key):
    v = FIX["vdos"][key]
    if v["kind"] == "debye":
        lo, hi, dens = oracle.vdos_debye(v["debye_temp"])
        return np.array([lo, hi]), dens, v["mass"]
    return np.array(v["egrid"]), np.array(v["density"]), v["mass"]


def test_regularise(oracle):
    for r in FIX["regularise"]:
        eg, dens, _ = _vdos(oracle, r["vdos"])
        lo, hi, d = oracle.vdos_regularise(eg, dens)
        assert d.size == r["npts"], r["vdos"]
        assert abs(lo * 1e3 / r["emin_mev"] - 1) < 1e-10 and abs(hi * 1e3 / r["emax_mev"] - 1) < 1e-10, r["vdos"]
        # the log prints 9 to 14 digits
        for x, y in zip(list(d[:3]) + list(d[-3:]), r["first3"] + r["last3"]):
            assert abs(x - y) <= 5e-9 * abs(y) + 1e-300, (r["vdos"], x, y)


def test_analyse(oracle):
    for r in FIX["analyse"]:
        eg, dens, mass = _vdos(oracle, r["vdos"])
        lo, hi, d = oracle.vdos_regularise(eg, dens)
        a = oracle.vdos_analyse(lo, hi, d, r["temperature"], mass)
        for key in ("gamma0", "integral", "msd", "teff"):
            e = r["expected"][key]
            assert abs(a[key] - e) <= 1e-10 * abs(e), (r["vdos"], r["temperature"], key, a[key], e)


def _printed_moments(x, y, nmax=7):
    dx = np.diff(x)
    integral = math.fsum(sorted(dx * (y[:-1] + y[1:]))) / 2
    y = y / integral
    dydx = np.diff(y) / dx
    a = y[:-1] - x[:-1] * dydx
    mean = math.fsum(sorted(list(a / 2 * x[1:] ** 2) + list(-a / 2 * x[:-1] ** 2) + list(dydx / 3 * x[1:] ** 3)
                            + list(-dydx / 3 * x[:-1] ** 3)))
    u = x - mean

    def cm(n):
        u1, u2 = u ** (n + 1) / (n + 1), u ** (n + 2) / (n + 2)
        return math.fsum(sorted(list(a * u1[1:]) + list(-a * u1[:-1]) + list(dydx * u2[1:]) + list(-dydx * u2[:-1])))

    rms = math.sqrt(cm(2))
    out = {"integral": integral, "mean": mean, "rms": rms}
    for k in range(3, nmax + 1):
        out[{3: "skewness", 4: "kurtosis"}.get(k, f"{k}th moment (std)")] = cm(k) / rms ** k
    return out


def test_gn_moments(oracle):
    seen = set()
    for r in FIX["gn"]:
        key = (r["vdos"], r["temperature"])
        if key in seen:  # the log repeats each case for vdoslux 0 and 3
            continue
        seen.add(key)
        eg, dens, mass = _vdos(oracle, r["vdos"])
        lo, hi, d = oracle.vdos_regularise(eg, dens)
        for order, expected in r["orders"].items():
            x, g = oracle.vdos_gn(lo, hi, d, r["temperature"], mass, int(order))
            got = _printed_moments(x, g)
            for name, e in expected.items():
                assert abs(got[name] - e) <= 6e-7 * abs(e), (r["vdos"], r["temperature"], order, name, got[name], e)


def test_kernel_shapes_and_errors(oracle):
    for r in FIX["kernel"]:
        lo, hi, dens = oracle.vdos_debye(r["debye_temp"])
        lo, hi, d = oracle.vdos_regularise(np.array([lo, hi]), dens)
        if "error" in r:
            with pytest.raises(RuntimeError) as exc:
                oracle.vdos_kernel(lo, hi, d, r["temperature"], r["mass"], vdoslux=r["vdoslux"],
                                   target_emax=r["target_emax"])
            assert str(exc.value) == r["error"]
            continue
        k = oracle.vdos_kernel(lo, hi, d, r["temperature"], r["mass"], vdoslux=r["vdoslux"],
                               target_emax=r["target_emax"])
        sf = oracle.sab_std_format(0, k["alpha"], k["beta"], k["sab"], beta_grid_optimised=True)
        assert (sf["alpha"].size, sf["beta"].size) == (r["nalpha"], r["nbeta"])
        assert sf["ntrimmed"] == r["discarded_edges"]
        assert k["suggested_emax"] == r["suggested_emax"]


@pytest.mark.parametrize("T", [10.0, 293.15, 1200.0])
def test_debye_vdos_msd(oracle, T):
    """a Debye VDOS sampled on a fine regular grid has the Debye model's MSD
    (to the accuracy of its linear interpolation)"""
    td, mass = 410.0, 26.98
    ed = KB * td
    n = 2001
    eg = np.linspace(ed / (n - 1) * 20, ed, n - 20)
    dens = (eg / ed) ** 2
    lo, hi, d = oracle.vdos_regularise(np.array([eg[0], eg[-1]]), dens)
    a = oracle.vdos_analyse(lo, hi, d, T, mass)
    ref = oracle.debye_msd(td, T, mass)
    assert abs(a["msd"] / ref - 1) < 1e-5, (a["msd"], ref)
    # effective temperature of a Debye spectrum: 3/8 T_D as T -> 0, T at high T
    if T == 10.0:
        assert abs(a["teff"] / (0.375 * td) - 1) < 1e-3
    if T == 1200.0:
        assert abs(a["teff"] / T - 1) < 0.02


def test_fill_independent_sum(oracle):
    """S(alpha, beta) of the kernel = sum_n x^n e^-x / n! kT G_n(beta kT),
    x = alpha * 2 m kT msd / hbar^2, over the orders the fill uses (blocks of
    at least 16 orders when there are more than 16, remainder dropped as the
    reference's block split does)"""
    lo, hi, dens = oracle.vdos_debye(400.0)
    lo, hi, d = oracle.vdos_regularise(np.array([lo, hi]), dens)
    T, mass = 300.0, 27.0
    k = oracle.vdos_kernel(lo, hi, d, T, mass, vdoslux=1)
    nord = k["norders"]
    used = nord if nord <= 16 else (nord // 16) * (nord // (nord // 16))
    kT = KB * T
    hbar = 4.135667662e-15 / (2 * math.pi)
    mn = 1.00866491588 * 931494095.17 / (299792458e10) ** 2
    a2x = (2.0 * mn / (hbar * hbar)) * kT * k["msd"]
    gs = [oracle.vdos_gn(lo, hi, d, T, mass, n) for n in range(1, used + 1)]
    rng = np.random.default_rng(1)
    alpha, beta, sab = k["alpha"], k["beta"], k["sab"]
    izero = int(np.argmin(np.abs(beta)))
    checked = 0
    for _ in range(300):
        j = int(rng.integers(0, izero + 1))
        i = int(rng.integers(0, alpha.size))
        x = alpha[i] * a2x
        e = beta[j] * kT
        terms = []
        for n, (eg, g) in enumerate(gs, start=1):
            if not (eg[0] <= e <= eg[-1]):
                continue
            gv = np.interp(e, eg, g)
            if gv > 0 and x > 0:
                terms.append(math.exp(n * math.log(x) - x - math.lgamma(n + 1)) * kT * gv)
        ref = math.fsum(terms)
        got = sab[j, i]
        if ref > 1e-280:
            assert abs(got - ref) <= 1e-9 * ref, (i, j, got, ref)
            checked += 1
    assert checked > 100
    # detailed balance on the mirrored rows
    for j in range(izero + 1, beta.size):
        jm = int(np.argmin(np.abs(beta + beta[j])))
        assert beta[jm] == -beta[j]
        ok = sab[jm] > 1e-300
        assert np.allclose(sab[j, ok], sab[jm, ok] * math.exp(-beta[j]), rtol=1e-13, atol=0)


def test_basic_log_aluminium(oracle):
    """tests/scripts/basic.log, aluminium of data/Al_sg225.ncmat at 293.15 K:
    the MSD from its VDOS into the Debye-Waller factors of the space-group-225
    reflections (dcutoff 0.1 Aa) gives the PowderBragg cross sections at
    1-4 Aa; 1/v absorption at 4 Aa; incoherent elastic plus the inelastic
    S(alpha,beta) expanded from the same VDOS (vdoslux 3, suggested Emax) at
    4 Aa gives the "Bkgd" value (6 printed digits each)"""
    b = FIX["basic"]
    eg, dens, mass = _vdos(oracle, b["vdos"])
    assert mass == b["mass"]
    T = b["temperature"]
    lo, hi, d = oracle.vdos_regularise(eg, dens)
    msd = oracle.vdos_analyse(lo, hi, d, T, mass)["msd"]
    a = b["a"]
    hkl, mult, dsp, fsq, _, _ = oracle.hkl_spacegroup_members(b["spacegroup"], (a, a, a, 90, 90, 90),
                                                               [(b["coh_sl"], msd, np.array(b["positions"]))],
                                                               b["dcutoff"])
    wl = np.array([1.0, 2.0, 3.0, 4.0])
    e = 0.081804209605330899 / wl ** 2
    pb = oracle.powderbragg(a ** 3 * len(b["positions"]), dsp, fsq * mult).xs(e)
    assert np.allclose(pb, b["powderbragg_1to4"], rtol=5e-6, atol=0)
    assert abs(oracle.absoov(b["sigma_abs"]).xs(e[3:])[0] / b["absorption_4"] - 1) < 5e-6
    k = oracle.vdos_kernel(lo, hi, d, T, mass, vdoslux=3)
    sf = oracle.sab_std_format(0, k["alpha"], k["beta"], k["sab"], beta_grid_optimised=True)
    sigma = 4 * math.pi * b["coh_sl"] ** 2 + b["sigma_inc"]
    inel = oracle.sab(sf["alpha"], sf["beta"], sf["sab"], T, mass, sigma, suggested_emax=k["suggested_emax"])
    bkgd = oracle.elinc([msd], [b["sigma_inc"]], [1.0]).xs(e[3:])[0] + inel.xs(e[3:])[0]
    assert abs(bkgd / b["bkgd_4"] - 1) < 5e-6, bkgd
    assert abs((bkgd + pb[3]) / b["bkgd_plus_bragg_4"] - 1) < 5e-6
