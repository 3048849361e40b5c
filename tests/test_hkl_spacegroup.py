"""hkl enumeration with a space group (NCrystal's calculateHKLPlanesWithSymEqRefl,
ncrystal_core/src/extd_utils/NCFillHKL.cc:545-709, equivalent reflections of
ncrystal_core/src/phys_utils/NCEqRefl.cc:25-150) in the CPU checker, pinned by

* the reference's own tables in tests/scripts/ncmat2hkl.log (committed as
  tests/golden/hkl_spacegroup.json): family count, hkl and multiplicity exact,
  d-spacings to the 10 printed digits, for cubic (225), monoclinic (15,
  beta = 122 degrees), orthorhombic (62), tetragonal (129), trigonal (167),
  hexagonal (186) and triclinic (Wollastonite, 2: the general lattice
  branch, NCLatticeUtils.cc:29-90 and 145-148, whose d-spacings and cell
  volume follow from the reference's lattice matrix as it stands) crystals;
  |F|^2 to the printed digits for Al, whose
  mean-squared displacement the log's effective Debye temperature fixes
  (the multi-element materials' per-element MSDs come from VDOS curves the
  log does not give, so their |F|^2 is not compared);
* brute force for every Laue class: with one atom per cell no reflection is
  extinct, so the families must partition the half-space of (h,k,l) with
  d >= dcutoff into orbits of equal d, each member listed once.
"""
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN

FIX = json.loads((GOLDEN / "hkl_spacegroup.json").read_text())["materials"]


def _species(oracle, m):
    single = len(m["species"]) == 1
    return [(s["csl_fm"] / 10.0,
             oracle.debye_msd(m["debye_temperature"], m["temperature"], m["average_mass"] if single else s["mass"]),
             s["positions"]) for s in m["species"]]


@pytest.mark.parametrize("name", sorted(FIX))
def test_reference_tables(oracle, name):
    m = FIX[name]
    hkl, mult, dsp, fsq, moff, mhkl = oracle.hkl_spacegroup_members(m["spacegroup"], m["cell"], _species(oracle, m),
                                                                     m["dcutoff"])
    head = np.array(m["table"])
    tail = np.array(m["tail"]).reshape(-1, 6)
    assert len(dsp) == len(head) + m["snipped"] + len(tail)
    idx = list(range(len(head))) + list(range(len(dsp) - len(tail), len(dsp)))
    rows = np.vstack([head, tail])
    assert np.array_equal(hkl[idx], rows[:, :3].astype(int))
    assert np.array_equal(mult[idx], rows[:, 3].astype(int))
    assert np.max(np.abs(dsp[idx] - rows[:, 4]) / rows[:, 4]) < 1e-9
    if len(m["species"]) == 1:
        assert np.max(np.abs(fsq[idx] - rows[:, 5]) / rows[:, 5]) < 1e-9
    # members: mult/2 distinct reflections per family, representative first
    assert np.array_equal(np.diff(moff), mult // 2)
    assert np.array_equal(mhkl[moff[:-1]], hkl)


def test_invalid_input_is_refused(oracle):
    with pytest.raises(RuntimeError, match="-4"):
        oracle.hkl_spacegroup_members(231, (4.0, 4.0, 4.0, 90, 90, 90), [(0.5, 0.001, [[0, 0, 0]])], 1.0)
    with pytest.raises(RuntimeError, match="-3"):   # angles that admit no cell
        oracle.hkl_spacegroup_members(2, (4.0, 4.0, 4.0, 60, 60, 179), [(0.5, 0.001, [[0, 0, 0]])], 1.0)


def test_general_lattice_matrix(oracle):
    """The general branch against what fixes it independently: the cell
    volume the reference prints for Wollastonite (ncmat2hkl.log: Vc
    373.76436342422 Aa^3) is |det| of its lattice matrix = (2 pi)^3 / |det| of
    the reciprocal one; reciprocal and direct lattice satisfy row_i(L) .
    column_j(M) = 2 pi delta_ij; a cell a hair away from the closed forms
    (outside their 1e-14 tolerance) takes the general branch and gives the
    closed forms' matrices to rounding.  (The reference's matrix has
    c (cos a - cos b cos g) / sin g in both off-diagonal places, so its second
    row is longer than b and, for example, a rhombohedral cell's (100), (010)
    and (001) planes get different d; the reference's tables follow from it
    and parity keeps it.)"""
    m = FIX["CaSiO3_sg2_Wollastonite.ncmat"]
    M = oracle.reci_lattice(m["cell"]).reshape(3, 3)
    assert abs((2 * math.pi) ** 3 / abs(np.linalg.det(M)) - 373.76436342422) < 1e-9
    L = 2 * math.pi * np.linalg.inv(M)
    assert np.allclose(L @ M, 2 * math.pi * np.eye(3), rtol=0, atol=1e-12)
    a, b, c, al, be, ga = m["cell"]
    assert abs(L[0, 0] - a) < 1e-12 and abs(L[0, 1]) < 1e-12 and abs(L[0, 2]) < 1e-12
    assert abs(np.linalg.norm(L[2]) - c) < 1e-12      # the third row keeps the length c
    for cell in [(3.1, 3.7, 4.3, 90, 90, 90), (3.2, 3.2, 4.4, 90, 90, 120), (3.1, 3.7, 4.3, 90, 104.5, 90)]:
        near = tuple(x + (1e-9 if i >= 3 else 0.0) for i, x in enumerate(cell))
        assert np.allclose(oracle.reci_lattice(near), oracle.reci_lattice(cell), rtol=0, atol=1e-9)


# (space group, cell) per Laue class; cells with the metric the class needs
LAUE = [
    (1, (3.1, 3.7, 4.3, 90, 90, 90)),
    (2, (3.1, 3.7, 4.3, 97.3, 104.1, 88.2)),
    (2, (3.9, 3.9, 3.9, 76.5, 76.5, 76.5)),
    (14, (3.1, 3.7, 4.3, 90, 104.5, 90)),
    (62, (3.1, 3.7, 4.3, 90, 90, 90)),
    (80, (3.3, 3.3, 4.1, 90, 90, 90)),
    (129, (3.3, 3.3, 4.1, 90, 90, 90)),
    (146, (3.2, 3.2, 4.4, 90, 90, 120)),
    (166, (3.2, 3.2, 4.4, 90, 90, 120)),
    (173, (3.2, 3.2, 4.4, 90, 90, 120)),
    (194, (3.2, 3.2, 4.4, 90, 90, 120)),
    (200, (3.6, 3.6, 3.6, 90, 90, 90)),
    (225, (3.6, 3.6, 3.6, 90, 90, 90)),
]
ORDER = {1: 1, 2: 1, 14: 2, 62: 4, 80: 4, 129: 8, 146: 3, 166: 6, 173: 6, 194: 12, 200: 12, 225: 24}


@pytest.mark.parametrize("sg,cell", LAUE)
def test_orbits_partition_the_reflections(oracle, sg, cell):
    dcut = 0.9
    hkl, mult, dsp, fsq, moff, mhkl = oracle.hkl_spacegroup_members(sg, cell, [(0.5, 1e-4, [[0.0, 0.0, 0.0]])],
                                                                     dcut, fsquarecut=1e-12)
    m = oracle.reci_lattice(cell).reshape(3, 3)
    # every member has its family's d, and no reflection is listed twice
    dmem = 2 * math.pi / np.linalg.norm(mhkl @ m.T, axis=1)
    fam = np.repeat(np.arange(len(dsp)), np.diff(moff))
    assert np.max(np.abs(dmem - dsp[fam]) / dsp[fam]) < 1e-12
    key = {tuple(x) for x in mhkl}
    assert len(key) == len(mhkl)
    # the (h,k,l) half-space (one of each +-pair) with d >= dcutoff, brute force
    n = 12
    g = np.array([(h, k, l) for h in range(-n, n + 1) for k in range(-n, n + 1) for l in range(-n, n + 1)
                  if (h, k, l) > (0, 0, 0)])
    dd = 2 * math.pi / np.linalg.norm(g @ m.T, axis=1)
    inside = {tuple(x) for x in g[dd >= dcut]}
    # members are stored as the form of (x, -x) whose first non-zero index is positive
    assert all(tuple(x) > (0, 0, 0) for x in mhkl)
    assert key == inside
    # orbit sizes divide the order of the Laue group (as +-pairs)
    assert all(ORDER[sg] % (k // 2) == 0 for k in mult)


@pytest.mark.parametrize("cell", [(3.1, 3.7, 4.3, 97.3, 104.1, 88.2), (3.9, 3.9, 3.9, 76.5, 76.5, 76.5)])
def test_cell_without_symmetry_matches_p1(oracle, cell):
    """the enumeration without symmetry (NCFillHKL.cc:230-450) on a general
    triclinic or rhombohedral cell against the space-group path with P1: with
    atoms at general positions no two reflections other than Friedel pairs
    share d and |F|^2, so both give the same families (d, |F|^2,
    multiplicity 2) and members; and the general entry on a 90/90/90 cell
    reproduces the orthogonal one exactly"""
    rng = np.random.default_rng(3)
    sp = [(0.3 + 0.2 * i, 0.002 * (i + 1), rng.uniform(0, 1, size=(2 + i, 3))) for i in range(3)]
    g = oracle.hkl_cell_members(cell, sp, 0.6, fsquarecut=1e-12)
    p = oracle.hkl_spacegroup_members(1, cell, sp, 0.6, fsquarecut=1e-12)
    assert len(g[2]) == len(p[2]) > 50
    og = np.lexsort((g[3], g[2]))
    op = np.lexsort((p[3], p[2]))
    assert np.max(np.abs(g[2][og] - p[2][op]) / p[2][op]) < 1e-12
    assert np.max(np.abs(g[3][og] - p[3][op]) / p[3][op]) < 1e-9
    assert np.all(g[1] == 2) and np.all(p[1] == 2)
    ortho = (3.1, 3.7, 4.3)
    a = oracle.hkl_cell_members(ortho + (90, 90, 90), sp, 0.6)
    b = oracle.hkl_orthogonal_members(*ortho, sp, 0.6)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
