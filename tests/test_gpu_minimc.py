"""MiniMC on the GPU (ncamd_minimc_sphere) against the reference's MiniMC
histograms of aluminium (p-value of Hist1D.check_compat at least 0.001, as
the reference's own tests require) and against the checker's restatement
(oracle/pymmc.py, itself pinned to the same histograms) with more neutrons,
category by category; weight conservation without absorption; a circular
beam wider than the sphere (misses tallied); the single-scattering limit;
every tally and roulette setting of the reference's tally and roulette
tests; argument errors."""
import numpy as np
import pytest

import mmc_common as M
import pymmc

pytestmark = pytest.mark.gpu

# about 30 statistical comparisons per run: a threshold of 1e-4 keeps the
# chance of any false alarm near 0.3%, while a 10% error in the material's
# temperature fails the energy tallies at p < 1e-28
P_MIN = 1e-4


@pytest.fixture(scope="module")
def amd():
    from ncrystal_amd import _capi, core
    if _capi.lib().ncamd_device_count() < 1:
        pytest.fail("no GPU visible to libncrystal_amd")
    return _capi, core


def _setup(core, T, wl, absorption, diam):
    scatter, ab, nd = M.library_al(T, absorption)
    ekin = M.wl2ekin(wl)
    r, z = M.sphere_setup(scatter.crossSectionIsotropic(np.array([ekin]))[0], ekin, diam)
    return scatter, ab, nd, ekin, r, z


@pytest.mark.parametrize("case", M.AL_CASES, ids=[c[0] for c in M.AL_CASES])
def test_matches_reference_histogram(amd, case):
    C, core = amd
    name, T, wl, absorption, diam = case
    scatter, ab, nd, ekin, r, z = _setup(core, T, wl, absorption, diam)
    res = core.minimc_sphere(scatter, ab, nd, r, ekin, 100000, pos=(0.0, 0.0, z), seed=5)
    ref = M.ref_hist(name)
    p = M.pvalue(res["total"], ref)
    print(f"{name}: integral {res['total']['content'].sum():.1f} (reference {ref['content'].sum():.1f}), "
          f"p-value {p:.3g}, {res['info']}")
    assert p >= P_MIN
    parts = sum(res[c]["content"] for c in core.MINIMC_CATEGORIES)
    assert np.allclose(parts, res["total"]["content"], rtol=1e-9, atol=1e-9)


@pytest.fixture(scope="module")
def al_by_temperature(amd):
    C, core = amd
    cache = {}

    def get(T):
        if T not in cache:
            cache[T] = M.library_al(T, True)
        return cache[T]
    return get


@pytest.mark.parametrize("case", M.TALLY_CASES, ids=[c[0] for c in M.TALLY_CASES])
def test_tallies_and_roulette_match_reference(amd, al_by_temperature, case):
    """every tally at the reference's binning and every roulette setting of
    the reference's tests (mmctally.py, mmcroul.py), 1e5 neutrons in 20
    independent batches (variance factor for correlated entries, sparse bins
    merged; the comparison of tests/test_minimc_checker.py)"""
    C, core = amd
    name, T, wl, tally, roul = case
    scatter, ab, nd = al_by_temperature(T)
    ekin = M.wl2ekin(wl)
    r, z = M.sphere_setup(scatter.crossSectionIsotropic(np.array([ekin]))[0], ekin, None)
    ref = M.ref_hist(name)

    def run(seed, n):
        return core.minimc_sphere(scatter, ab, nd, r, ekin, n, pos=(0.0, 0.0, z), seed=seed, tally=tally,
                                  nbins=ref["nbins"], xrange=(ref["xmin"], ref["xmax"]), roulette=roul)
    h, phi = M.batched(run, 100000)
    p = M.pvalue_phi(h, ref, phi)
    print(f"{name}: integral {h['content'].sum():.1f} (reference {ref['content'].sum():.1f}), phi {np.mean(phi):.2f}-{np.max(phi):.2f}, "
          f"p {p:.3g}")
    assert p >= P_MIN


def test_matches_checker_by_category(amd, oracle):
    C, core = amd
    scatter, ab, nd, ekin, r, z = _setup(core, 293.15, 4.0, True, None)
    n = 1000000
    g = core.minimc_sphere(scatter, ab, nd, r, ekin, n, pos=(0.0, 0.0, z), seed=7)
    cs, cab, _ = M.checker_al(oracle, 293.15, True)
    o = pymmc.minimc_sphere(cs, cab, nd, r, ekin, n, pos=(0.0, 0.0, z), seed=8, nthreads=8)
    for key in ("total",) + tuple(core.MINIMC_CATEGORIES):
        p = M.pvalue(g[key], o[key])
        print(f"{key}: GPU {g[key]['content'].sum():.1f} checker {o[key]['content'].sum():.1f} p-value {p:.3g}")
        assert p >= P_MIN, key
        # the weight in each category agrees within 5 standard errors
        sg, so = g[key]["content"].sum(), o[key]["content"].sum()
        err = np.sqrt(g[key]["errorsq"].sum() + o[key]["errorsq"].sum())
        assert abs(sg - so) <= 5 * err + 1e-9, key


def test_weight_conserved_without_absorption(amd):
    C, core = amd
    scatter, ab, nd, ekin, r, z = _setup(core, 293.15, 2.5, False, 0.2)
    n = 200000
    res = core.minimc_sphere(scatter, None, nd, r, ekin, n, pos=(0.0, 0.0, z), seed=3)
    tot = res["total"]["content"].sum() + res["total"]["underflow"] + res["total"]["overflow"]
    print("integral / n", tot / n)
    assert abs(tot / n - 1.0) < 5e-3


def test_circular_beam_with_misses(amd, oracle):
    """a uniform beam of radius 1.5 r: a 1 - 1/1.5^2 share misses the sphere and
    is tallied unscattered at theta = 0"""
    C, core = amd
    scatter, ab, nd, ekin, r, z = _setup(core, 293.15, 4.0, True, None)
    n = 400000
    kw = dict(source="circular", pos=(0.0, 0.0, -1.5 * r), beam_radius=1.5 * r)
    g = core.minimc_sphere(scatter, ab, nd, r, ekin, n, seed=9, **kw)
    cs, cab, _ = M.checker_al(oracle, 293.15, True)
    o = pymmc.minimc_sphere(cs, cab, nd, r, ekin, n, seed=10, nthreads=8, **kw)
    p = M.pvalue(g["total"], o["total"])
    miss = n * (1 - 1 / 1.5 ** 2)
    print("p-value", p, "first bin", g["total"]["content"][0], "misses expected", miss)
    assert p >= P_MIN
    assert g["total"]["content"][0] > miss
    gi = core.minimc_sphere(scatter, ab, nd, r, ekin, n, seed=9, ignoremiss=True, **kw)
    assert abs(g["total"]["content"][0] - gi["total"]["content"][0] - miss) < 5 * np.sqrt(miss)


def test_single_scattering_limit(amd, oracle):
    """nscatlimit = 1 and the nscat tally: no neutron scatters twice"""
    C, core = amd
    scatter, ab, nd, ekin, r, z = _setup(core, 293.15, 4.0, True, None)
    kw = dict(pos=(0.0, 0.0, z), tally="nscat", nbins=22, xrange=(-1.5, 20.5), nscatlimit=1)
    g = core.minimc_sphere(scatter, ab, nd, r, ekin, 200000, seed=2, **kw)
    assert np.all(g["total"]["content"][3:] == 0.0) and g["total"]["content"][2] > 0
    cs, cab, _ = M.checker_al(oracle, 293.15, True)
    o = pymmc.minimc_sphere(cs, cab, nd, r, ekin, 200000, seed=4, nthreads=8, **kw)
    p = M.pvalue(g["total"], o["total"])
    print("nscat p-value", p)
    assert p >= P_MIN


def test_argument_errors(amd):
    C, core = amd
    scatter, ab, nd, ekin, r, z = _setup(core, 293.15, 4.0, True, None)
    with pytest.raises(C.NCBadInput):
        core.minimc_sphere(scatter, ab, nd, -1.0, ekin, 10)
    with pytest.raises(C.NCBadInput):
        core.minimc_sphere(scatter, ab, nd, r, ekin, 10, nbins=1000)
    with pytest.raises(C.NCBadInput):
        core.minimc_sphere(ab, scatter, nd, r, ekin, 10)
