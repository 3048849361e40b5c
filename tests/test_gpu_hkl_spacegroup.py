"""Space-group hkl enumeration through the C ABI on the GPU (|F|^2 of every
symmetry representative on the device) against the CPU checker, which
tests/test_hkl_spacegroup.py pins to the reference's own tables: families,
members, multiplicities and d-spacings identical, |F|^2 within a few ulp (the
device exp), on the reference materials (the triclinic one included) and one
crystal per Laue class (two general triclinic cells among them); a
mosaic single crystal of hexagonal ZnO (sg 186) built from those families
against the checker."""
import math

import numpy as np
import pytest

from ncrystal_amd import synth
from test_hkl_spacegroup import FIX, LAUE, _species
from conftest import assert_streams_match

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return core


def _compare(g, o):
    ghkl, gmult, gdsp, gfsq, gmoff, gmhkl = g
    ohkl, omult, odsp, ofsq, omoff, omhkl = o
    assert np.array_equal(ghkl, ohkl) and np.array_equal(gmult, omult)
    assert np.array_equal(gmoff, omoff) and np.array_equal(gmhkl, omhkl)
    assert np.array_equal(gdsp, odsp)
    assert np.max(np.abs(gfsq - ofsq) / ofsq) < 1e-14


@pytest.mark.parametrize("name", sorted(FIX))
def test_reference_materials(core, oracle, name):
    m = FIX[name]
    sp = _species(oracle, m)
    g = core.hkl_spacegroup_members(m["spacegroup"], m["cell"], sp, m["dcutoff"])
    o = oracle.hkl_spacegroup_members(m["spacegroup"], m["cell"], sp, m["dcutoff"])
    _compare(g, o)
    assert np.array_equal(core.reci_lattice(m["cell"]), oracle.reci_lattice(m["cell"]))


@pytest.mark.parametrize("sg,cell", LAUE)
def test_laue_classes(core, oracle, sg, cell):
    # three species at general positions: |F|^2 differs between reflections
    rng = np.random.default_rng(sg)
    sp = [(0.2 + 0.3 * i, 0.002 * (i + 1), rng.uniform(0, 1, size=(2 + i, 3))) for i in range(3)]
    g = core.hkl_spacegroup_members(sg, cell, sp, 0.5)
    o = oracle.hkl_spacegroup_members(sg, cell, sp, 0.5)
    assert len(g[0]) > 50
    _compare(g, o)


def test_bad_inputs(core):
    from ncrystal_amd import _capi as C
    sp = [(0.5, 0.001, [[0.0, 0.0, 0.0]])]
    with pytest.raises(C.NCBadInput):
        core.hkl_spacegroup_members(2, (4.0, 4.0, 4.0, 60, 60, 179), sp, 1.0)   # no such cell
    with pytest.raises(C.NCBadInput):
        core.hkl_spacegroup_members(0, (4.0, 4.0, 4.0, 90, 90, 90), sp, 1.0)
    with pytest.raises(C.NCBadInput):
        core.reci_lattice((4.0, 4.0, 4.0, 0.0, 90, 90))


def test_scbragg_hexagonal_zno(core, oracle):
    m = FIX["ZnO_sg186_ZincOxide.ncmat"]
    sp = _species(oracle, m)
    hkl, mult, dsp, fsq, moff, mh = core.hkl_spacegroup_members(m["spacegroup"], m["cell"], sp, 0.4)
    L = core.reci_lattice(m["cell"])
    a, b, c = m["cell"][:3]
    v0n = a * b * c * math.sin(math.radians(120.0)) * sum(len(s["positions"]) for s in m["species"])
    R = synth.random_rotation(7)
    mos = 20 * math.pi / (180 * 60)
    g = core.create_scbragg(v0n, dsp, fsq, moff, mh, L, R, mos)
    o = oracle.scbragg(a, b, c, v0n, dsp, fsq, moff, mh, R, mos, reci=L)
    n = 20000
    wl = synth.wavelengths_uniform(n, 0.8, 6.0, seed=2)
    e = 0.081804209605330899 / wl ** 2
    dirs = synth.isotropic_dirs(n, seed=5)
    xg = g.crossSection(e, dirs)
    xo = o.xs_dir(e, dirs, nthreads=8)
    den = np.maximum(np.abs(xg), np.abs(xo))
    assert np.all(np.where(den > 0, np.abs(xg - xo) / np.where(den > 0, den, 1.0), 0.0) <= 1e-12)
    assert np.mean(xg > 0) > 0.05
    g.setRandomSeed(3)
    ge, gd = g.sampleScatter(e, dirs)
    oe, od, nerr = o.sample_dir(e, dirs, seed=3, nthreads=8)
    assert nerr == 0 and np.array_equal(ge, e)
    assert_streams_match(np.max(np.abs(gd - od), axis=1) > 1e-9, "space-group mosaic crystal")


@pytest.mark.parametrize("cell", [(3.1, 3.7, 4.3, 97.3, 104.1, 88.2), (3.9, 3.9, 3.9, 76.5, 76.5, 76.5),
                                  (3.2, 3.2, 4.4, 90, 90, 120)])
def test_cell_without_symmetry(core, oracle, cell):
    """the enumeration without symmetry on general cells (and a hexagonal
    one) through ncamd_hkl_cell_members against the checker: same families
    and members, d and |F|^2 to rounding"""
    rng = np.random.default_rng(3)
    sp = [(0.3 + 0.2 * i, 0.002 * (i + 1), rng.uniform(0, 1, size=(2 + i, 3))) for i in range(3)]
    g = core.hkl_cell_members(cell, sp, 0.6, fsquarecut=1e-12)
    o = oracle.hkl_cell_members(cell, sp, 0.6, fsquarecut=1e-12)
    assert len(g[2]) > 50
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1])
    assert np.array_equal(g[4], o[4]) and np.array_equal(g[5], o[5])
    assert np.max(np.abs(g[2] - o[2]) / o[2]) < 1e-14
    assert np.max(np.abs(g[3] - o[3]) / o[3]) < 1e-13
