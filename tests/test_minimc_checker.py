"""The checker's MiniMC (oracle/pymmc.py) pinned to the reference's MiniMC
histograms of aluminium (tests/scripts/mmc_al.py and mmc_alnoabs.py: pencil
beam into a sphere, 1e5 neutrons, theta tally, p-value of Hist1D.check_compat
at least 0.001 as in NCTestUtils/minimc_ref.py), and its geometry and
weight bookkeeping checked against closed forms."""
import numpy as np
import pytest

import mmc_common as M
import pymmc


@pytest.mark.parametrize("case", M.AL_CASES, ids=[c[0] for c in M.AL_CASES])
def test_checker_matches_reference_histogram(oracle, case):
    name, T, wl, absorption, diam = case
    scatter, ab, nd = M.checker_al(oracle, T, absorption)
    ekin = M.wl2ekin(wl)
    r, z = M.sphere_setup(scatter.xs(np.array([ekin]))[0], ekin, diam)
    res = pymmc.minimc_sphere(scatter, ab, nd, r, ekin, 100000, pos=(0.0, 0.0, z), seed=11, nthreads=8)
    ref = M.ref_hist(name)
    p = M.pvalue(res["total"], ref)
    tot = res["total"]["content"].sum()
    print(f"{name}: sphere radius {r:.4f} m, integral {tot:.1f} (reference {ref['content'].sum():.1f}), "
          f"p-value {p:.3g}, steps {res['info']['max_steps']}")
    assert p >= 1e-4  # many comparisons per run (family-wise rate near 0.3%)
    # the breakdown adds up to the total
    parts = sum(res[c]["content"] for c in pymmc.CATEGORIES)
    assert np.allclose(parts, res["total"]["content"], rtol=1e-12, atol=1e-12)


def test_sphere_helpers_closed_forms():
    r2 = 4.0
    p = np.array([[0.0, 0.0, -2.0], [0.0, 0.0, -3.0], [3.0, 0.0, -3.0], [0.0, 0.0, 0.0], [0.0, 0.0, 2.0],
                  [1.0, 1.0, -5.0]])
    u = np.array([[0.0, 0.0, 1.0]] * 6)
    ent = pymmc._sphere_entry(r2, p, u)
    # on the surface moving in: 0; outside on the axis: 1; off to the side: miss;
    # inside: 0; on the surface moving out: miss; off-axis chord: 5 - sqrt(2)
    assert np.allclose(ent, [0.0, 1.0, -1.0, 0.0, -1.0, 5.0 - np.sqrt(2.0)])
    ex = pymmc._sphere_exit(r2, np.array([[0.0, 0.0, -2.0], [0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), u[:3])
    assert np.allclose(ex, [4.0, 2.0, np.sqrt(3.0)])


def test_transmission_without_scattering(oracle):
    """absorption only: every neutron is tallied once, unscattered, with
    weight exp(-Sigma_a 2r); nothing scatters"""
    ab = oracle.absoov(2.0)
    ekin = 0.0253
    nd, r = 0.05, 0.01
    res = pymmc.minimc_sphere(None, ab, nd, r, ekin, 1000, pos=(0.0, 0.0, -r), seed=1)
    w = np.exp(-100.0 * nd * ab.xs(np.array([ekin]))[0] * 2 * r)
    assert res["info"]["max_steps"] == 0
    assert abs(res["total"]["content"][0] - 1000 * w) <= 1e-9 * 1000
    assert abs(res["NOSCAT"]["content"].sum() - 1000 * w) <= 1e-9 * 1000


@pytest.mark.parametrize("case", M.TALLY_CASES, ids=[c[0] for c in M.TALLY_CASES])
def test_checker_tallies_and_roulette(oracle, case):
    """every tally (at the reference's binning) and roulette setting; the
    chi-square uses variances corrected for the correlation of a history's
    entries (factor phi from 20 independent batches), which the reference's
    sum of squared weights leaves out: two checker runs differ at p ~ 1e-4
    on the energy tallies without it"""
    name, T, wl, tally, roul = case
    scatter, ab, nd = M.checker_al(oracle, T, True)
    ekin = M.wl2ekin(wl)
    r, z = M.sphere_setup(scatter.xs(np.array([ekin]))[0], ekin, None)
    ref = M.ref_hist(name)

    def run(seed, n):
        return pymmc.minimc_sphere(scatter, ab, nd, r, ekin, n, pos=(0.0, 0.0, z), seed=seed, sample_seed=seed,
                                   nthreads=8, tally=tally, nbins=ref["nbins"], xrange=(ref["xmin"], ref["xmax"]),
                                   roulette=roul)
    h, phi = M.batched(run, 100000)
    p = M.pvalue_phi(h, ref, phi)
    print(f"{name}: integral {h['content'].sum():.1f} (reference {ref['content'].sum():.1f}), phi {np.mean(phi):.2f}-{np.max(phi):.2f}, "
          f"p {p:.3g}")
    assert p >= 1e-4  # many comparisons per run (family-wise rate near 0.3%)
