"""Compositions built as ProcComposition::addComponent builds them
(ncrystal_core/src/interfaces/NCProcImpl.cc:263-338) in the CPU checker,
pinned by what merging must preserve:

* PowderBragg + PowderBragg (NCPowderBragg.cc:218-331): the merged table is
  the table of one PowderBragg process built from the union of both plane
  lists with |F|^2 m rescaled by scale / v0n (same 2dE points, same
  cumulative sums), and its cross section is s1 xs1 + s2 xs2;
* ElInc + ElInc (NCElIncXS.cc:204-233): a hand-worked element list (sorted,
  zero strengths dropped, msd equal within 1e-6 summed) and the cross
  section s1 xs1 + s2 xs2 when the msd values coincide;
* AbsOOV + AbsOOV (NCAbsOOV.cc:60-70): c = s1 c1 + s2 c2;
* a process given twice adds its scales, a nested composition is flattened
  with multiplied scales, zero scales and empty PowderBragg are dropped, and
  a merged component keeps the position of the first one.
"""
import numpy as np
import pytest

from ncrystal_amd import synth


def _pb(oracle, cr):
    sp = [(s.csl, oracle.debye_msd(s.debye_temp, cr.temperature, s.mass), s.positions) for s in cr.species]
    hkl, mult, dsp, fsq, _, _ = oracle.hkl_orthogonal_members(cr.a, cr.b, cr.c, sp, cr.dcutoff)
    v0n = cr.volume * cr.n_atoms
    return v0n, dsp, fsq * mult


@pytest.mark.parametrize("second", ["fcc_double", "ortho"])
def test_powderbragg_merge_is_the_union(oracle, second):
    a = synth.fcc_powder(dcutoff=0.3)
    # a lattice twice as large repeats every d-spacing of the first one (to
    # within rounding): those points merge
    b = synth.fcc_powder(a=8.1, dcutoff=0.3) if second == "fcc_double" else synth.ortho_random(seed=3, dcutoff=0.3)
    va, da, fa = _pb(oracle, a)
    vb, db, fb = _pb(oracle, b)
    sa, sb = 0.37, 1.9
    pa, pb = oracle.powderbragg(va, da, fa), oracle.powderbragg(vb, db, fb)
    comp = oracle.composition([pa, pb], [sa, sb])
    parts = oracle.components(comp)
    assert [(t, s) for t, s, _ in parts] == [("PowderBragg", 1.0)]
    e2d, fdm = parts[0][2]
    union = oracle.powderbragg(1.0, np.concatenate([da, db]), np.concatenate([fa * (sa / va), fb * (sb / vb)]))
    ue, uf = oracle.pb_tables(union)
    na, nb = len(oracle.pb_tables(pa)[0]), len(oracle.pb_tables(pb)[0])
    assert len(ue) == len(e2d) and len(e2d) <= na + nb
    if second == "fcc_double":
        assert len(e2d) == nb < na + nb
    assert np.max(np.abs(e2d - ue) / ue) < 1e-13
    assert np.max(np.abs(fdm - uf) / uf[-1]) < 1e-13
    assert np.all(np.diff(e2d) > 0)
    e = synth.energies_logu(20000, 1e-3, 2.0, seed=5)
    ref = sa * pa.xs(e) + sb * pb.xs(e)
    x = comp.xs(e)
    assert np.max(np.abs(x - ref) / np.maximum(ref, 1e-300)) < 1e-12
    assert np.array_equal(x == 0, ref == 0)


def test_powderbragg_merge_fails_a_wrong_scale(oracle):
    """the union pin above sees a scale applied to the wrong list"""
    a, b = synth.fcc_powder(dcutoff=0.3), synth.ortho_random(seed=3, dcutoff=0.3)
    va, da, fa = _pb(oracle, a)
    vb, db, fb = _pb(oracle, b)
    comp = oracle.composition([oracle.powderbragg(va, da, fa), oracle.powderbragg(vb, db, fb)], [0.37, 1.9])
    _, fdm = oracle.components(comp)[0][2]
    wrong = oracle.powderbragg(1.0, np.concatenate([da, db]), np.concatenate([fa * (1.9 / va), fb * (0.37 / vb)]))
    assert np.max(np.abs(fdm - oracle.pb_tables(wrong)[1])) > 1e-3 * fdm[-1]


def test_elinc_merge_by_hand(oracle):
    a = oracle.elinc([0.02, 0.01, 0.05], [2.0, 1.0, 0.0], [1.0, 1.0, 1.0])
    b = oracle.elinc([0.0100005, 0.03, 0.012], [0.5, 4.0, 1.5], [1.0, 1.0, 1.0])
    comp = oracle.composition([a, b], [0.3, 0.7])
    parts = oracle.components(comp)
    assert [(t, s) for t, s, _ in parts] == [("ElInc", 1.0)]
    msd, bixs = parts[0][2]
    # sorted by msd; 0.0100005 is within 1e-6 of 0.01 and merges into it,
    # 0.012 does not; the zero-strength element is dropped
    assert np.array_equal(msd, [0.01, 0.012, 0.02, 0.03])
    assert np.allclose(bixs, [0.3 + 0.35, 1.05, 0.6, 2.8], rtol=1e-15, atol=0)


def test_elinc_merge_xs(oracle):
    a = oracle.elinc([0.011, 0.02], [0.4, 5.0], [0.25, 0.75])
    b = oracle.elinc([0.02, 0.007], [3.0, 0.9], [0.5, 0.5])
    comp = oracle.composition([a, b], [0.6, 0.4])
    assert len(oracle.components(comp)[0][2][0]) == 3
    e = synth.energies_logu(5000, 1e-5, 10.0, seed=2)
    ref = 0.6 * a.xs(e) + 0.4 * b.xs(e)
    assert np.max(np.abs(comp.xs(e) - ref) / ref) < 1e-14


def test_absoov_merge(oracle):
    a, b = oracle.absoov(0.231), oracle.absoov(2.56)
    comp = oracle.composition([a, b], [0.25, 0.75])
    assert [t for t, _, _ in oracle.components(comp)] == ["AbsOOV"]
    e = np.geomspace(1e-5, 1.0, 100)
    assert np.allclose(comp.xs(e), 0.25 * a.xs(e) + 0.75 * b.xs(e), rtol=1e-15, atol=0)


def test_component_bookkeeping(oracle):
    cr = synth.fcc_powder(dcutoff=0.4)
    v, d, f = _pb(oracle, cr)
    pb1 = oracle.powderbragg(v, d, f)
    pb2 = oracle.powderbragg(v * 2.0, d, f)
    empty = oracle.powderbragg(v, np.zeros(0), np.zeros(0))
    ei = oracle.elinc([0.01], [0.5], [1.0])
    t = synth.small_sab_table(mass_amu=26.98)
    sab = oracle.sab(t["alpha"], t["beta"], t["sab"], t["temperature"], t["mass"], t["sigma_bound"])
    # the same process twice adds its scales; zero scale and empty Bragg drop out
    c1 = oracle.composition([ei, sab, ei, empty, pb1], [1.0, 2.0, 0.5, 1.0, 0.0])
    assert [(t_, s) for t_, s, _ in oracle.components(c1)] == [("ElInc", 1.5), ("SAB", 2.0)]
    # a nested composition is flattened with multiplied scales; the merged
    # Bragg component stays where the first Bragg component was
    inner = oracle.composition([pb1, sab], [1.0, 1.0])
    c2 = oracle.composition([inner, ei, pb2], [0.5, 1.0, 3.0])
    assert [(t_, s) for t_, s, _ in oracle.components(c2)] == [("PowderBragg", 1.0), ("SAB", 0.5), ("ElInc", 1.0)]
    e = synth.energies_logu(3000, 1e-4, 1.0, seed=1)
    ref = 0.5 * pb1.xs(e) + 0.5 * sab.xs(e) + ei.xs(e) + 3.0 * pb2.xs(e)
    assert np.max(np.abs(c2.xs(e) - ref) / ref) < 1e-12
