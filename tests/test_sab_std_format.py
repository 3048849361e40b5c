"""The checker's transformKernelToStdFormat (ncrystal_core/src/sab/NCSABUtils.cc:
134-325) pinned by what each step fixes:

* trimming: exactly the all-zero upper alpha columns, all-zero rows at
  beta > 0 from the top and at beta < 0 from the bottom are removed, nothing
  is removed from a table that is zero everywhere;
* symmetric tables: the beta grid is mirrored with one zero row, and
  S(alpha, -beta) = S(alpha, beta) before unscaling;
* unscaling: S = S_scaled exp(-beta/2), so the unscaled output of a
  symmetric table obeys detailed balance S(alpha, -beta) = e^beta
  S(alpha, beta); where exp(-beta/2) overflows the product is formed in log
  space (and an unrepresentable result is an error);
* thickening: grids of fewer than 500 rows get floor(500 / n) rows between
  neighbours, on the uniform subdivision, linear in beta column by column;
  the original rows are kept exactly; an optimised grid is left alone;
* the input checks of validateScatKnlData (NCScatKnlData.cc:26-80).
CPU only."""
import math

import numpy as np
import pytest


def _table(na=12, nb=40, seed=1, bmin=-8.0, bmax=8.0):
    rng = np.random.default_rng(seed)
    a = np.geomspace(0.01, 20.0, na)
    b = np.linspace(bmin, bmax, nb)
    s = rng.uniform(0.1, 2.0, size=(nb, na))
    return a, b, s


def test_thickening(oracle):
    a, b, s = _table(nb=60)
    r = oracle.sab_std_format(0, a, b, s)
    nextra = 500 // 60
    assert r["beta"].size == 1 + 59 * (1 + nextra) and r["ntrimmed"] == 0
    assert np.array_equal(r["alpha"], a)
    assert np.array_equal(r["beta"][::nextra + 1], b)
    assert np.array_equal(r["sab"][::nextra + 1], s)
    fine = np.concatenate([np.linspace(b[j], b[j + 1], nextra + 2)[:-1] for j in range(59)] + [b[-1:]])
    assert np.max(np.abs(r["beta"] - fine)) < 1e-14
    ref = np.stack([np.interp(r["beta"], b, s[:, i]) for i in range(a.size)], axis=1)
    assert np.max(np.abs(r["sab"] - ref) / ref) < 1e-14
    # an optimised grid, or one of 500 rows or more, is left alone
    r2 = oracle.sab_std_format(0, a, b, s, beta_grid_optimised=True)
    assert np.array_equal(r2["beta"], b) and np.array_equal(r2["sab"], s)
    a3, b3, s3 = _table(na=6, nb=500, seed=2)
    assert np.array_equal(oracle.sab_std_format(0, a3, b3, s3)["sab"], s3)


def test_trim_zero_edges(oracle):
    a, b, s = _table(na=10, nb=600)
    s[:, -3:] = 0.0   # three zero alpha columns at the top
    s[-4:, :] = 0.0   # four zero rows at beta > 0
    s[:2, :] = 0.0    # two zero rows at beta < 0
    s[300, :] = 0.0   # an inner zero row stays
    r = oracle.sab_std_format(0, a, b, s)
    assert r["ntrimmed"] == 9
    assert np.array_equal(r["alpha"], a[:-3]) and np.array_equal(r["beta"], b[2:-4])
    assert np.array_equal(r["sab"], s[2:-4, :-3])
    z = np.zeros_like(s)
    r = oracle.sab_std_format(0, a, b, z)
    assert r["ntrimmed"] == 0 and r["sab"].shape == z.shape


def test_scaled_and_symmetric(oracle):
    a, _, s = _table(na=8, nb=520)
    bh = np.linspace(0.0, 30.0, 520)
    r = oracle.sab_std_format(2, a, bh, s)
    full = np.concatenate([-bh[::-1], bh[1:]])
    assert np.array_equal(r["beta"], full) and not np.signbit(r["beta"][519])
    sym = np.concatenate([s[::-1], s[1:]])
    ref = sym * np.exp(-0.5 * full)[:, None]
    assert np.max(np.abs(r["sab"] - ref) / ref) < 1e-15
    # detailed balance of the unscaled table
    k = np.arange(1, 520)
    lhs, rhs = r["sab"][519 - k], r["sab"][519 + k] * np.exp(bh[k])[:, None]
    assert np.max(np.abs(lhs - rhs) / lhs) < 1e-13
    # scaled, non-symmetric, with rows where exp(-beta/2) overflows
    b = np.concatenate([[-1800.0, -1500.0], np.linspace(-10, 10, 600)])
    s2 = np.abs(_table(na=8, nb=602, seed=3)[2])
    s2[0] = 1e-320
    s2[1] = 1e-300
    r = oracle.sab_std_format(1, a, b, s2)
    ref = np.exp(-0.5 * b[2:])[:, None] * s2[2:]
    assert np.max(np.abs(r["sab"][2:] - ref) / ref) < 1e-15
    with np.errstate(over="ignore"):
        exact = np.exp(-0.5 * b[1] + np.log(s2[1]))
    assert np.max(np.abs(r["sab"][1] - exact) / exact) < 1e-14
    s2[0] = 1e-5  # exp(900 + log 1e-5) is not representable
    with pytest.raises(ValueError):
        oracle.sab_std_format(1, a, b, s2)


@pytest.mark.parametrize("bad", ["short", "unsorted", "alpha0", "sym_start", "neg_start", "negative_s", "nan"])
def test_input_checks(oracle, bad):
    a, b, s = _table(nb=20)
    kt = 0
    if bad == "short":
        a = a[:4]
        s = s[:, :4]
    elif bad == "unsorted":
        b = b.copy()
        b[5] = b[4]
    elif bad == "alpha0":
        a = np.concatenate([[0.0], a[1:]])
    elif bad == "sym_start":
        kt = 2
    elif bad == "neg_start":
        b = b - b[0] + 0.5
    elif bad == "negative_s":
        s = s.copy()
        s[3, 3] = -1.0
    elif bad == "nan":
        s = s.copy()
        s[3, 3] = math.nan
    with pytest.raises(ValueError):
        oracle.sab_std_format(kt, a, b, s)
